"""The C-ABI library loads on a machine without a GPU and exports every
symbol include/samhip.h declares (no compute calls here)."""
import ctypes
import os
import re

import pytest

from sam_amd import samhip as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "samhip.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(sam_[a-z0-9_]+)\s*\(", txt)))


def test_header_symbols_listed():
    assert set(declared_symbols()) == set(S.EXPORTS)


def test_library_exports_all_symbols():
    L = S.load()
    for name in declared_symbols():
        assert hasattr(L, name), name
        assert isinstance(getattr(L, name), ctypes._CFuncPtr)


def test_create_rejects_bad_params_before_touching_a_device():
    """parameter validation mirrors check_inputs.f's error classes and runs
    before any device query, so it is testable on a CPU host"""
    L = S.load()
    h = ctypes.c_void_p()
    p = S.make_params(48)                          # not a power of two
    assert L.sam_create(ctypes.byref(p), ctypes.byref(h)) == S.SAM_ERR_UNSUPPORTED
    p = S.make_params(32, n_dealias=3)              # n_dealias=3 has no case in init.f
    assert L.sam_create(ctypes.byref(p), ctypes.byref(h)) == S.SAM_ERR_VALUE
    p = S.make_params(32)
    p.iubc_z = 1                                     # sine/cosine z: out of scope
    assert L.sam_create(ctypes.byref(p), ctypes.byref(h)) == S.SAM_ERR_UNSUPPORTED
    p = S.make_params(32, i_cfl=0, dt0=0.0)          # dt0 required when i_cfl=0
    assert L.sam_create(ctypes.byref(p), ctypes.byref(h)) == S.SAM_ERR_REQUIRED
    assert b"dt0" in L.sam_last_error()
