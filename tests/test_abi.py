"""CPU-side checks of the boundary: libpmx.so loads, exports exactly what include/pmx.h declares,
and the ctypes signature table covers every declaration."""

from __future__ import annotations

import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(ROOT, "include", "pmx.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pmx_[a-z0-9_]+)\s*\(", text)))


def test_header_lists_entry_points():
    names = header_functions()
    assert "pmx_step" in names and "pmx_forces" in names and len(names) >= 20


def test_library_exports_every_declared_symbol(pkg):
    lib = pkg._lib.load()
    for name in header_functions():
        assert hasattr(lib, name), name


def test_signature_table_matches_header(pkg):
    assert sorted(pkg._lib.SIGNATURES) == header_functions()


def test_host_only_calls(pkg):
    lib = pkg._lib.load()
    assert lib.pmx_abi_version() == 1
    for n in (8, 16, 32, 64, 96, 128, 192, 256, 512, 1024):
        assert lib.pmx_fft_supported(n) == 1
    for n in (0, 7, 100, 2048):
        assert lib.pmx_fft_supported(n) == 0


def test_null_plan_is_rejected_without_device(pkg):
    lib = pkg._lib.load()
    rc = lib.pmx_plan_destroy(None)
    assert rc == 0
    h = ctypes.c_void_p()
    rc = lib.pmx_plan_create(ctypes.byref(h), 7, 8, 8, 0)
    assert rc == 2 and b"unsupported" in lib.pmx_last_error()
