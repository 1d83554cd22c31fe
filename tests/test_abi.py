"""CPU-side checks of the C ABI: the library builds, loads and exports every symbol
include/psx.h declares; host-side argument validation (no GPU compute)."""
import os
import re
import subprocess

import pytest

from pyspectral_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "psx.h")).read()
    return sorted(set(re.findall(r"\b(psx_[a-z_]+)\s*\(", txt)))


def test_header_and_binding_agree():
    assert _declared() == sorted(_lib.EXPORTS)


def test_library_exports_every_declared_symbol():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libpsx.so not built")
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\bT (psx_[a-z_]+)\b", out))
    missing = set(_declared()) - exported
    assert not missing, missing
    lib = _lib.load()
    for name in _lib.EXPORTS:
        assert hasattr(lib, name)


def test_null_context_is_rejected():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libpsx.so not built")
    lib = _lib.load()
    assert lib.psx_destroy(None) == -1
    assert lib.psx_step(None, None, None, None, 0.0) == -1
    assert lib.psx_last_error(None) == b"null context"


def test_params_struct_layout():
    import ctypes as C
    # 12 doubles (nu, L1-3, dx-dz, cfl, Cs, Ct, tau_a, tau_b), int kc, int rotate, 3 doubles
    assert C.sizeof(_lib.psx_params) == 12 * 8 + 2 * 4 + 4 * 8
