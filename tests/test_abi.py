"""CPU-side checks of the C ABI: the library exists, loads, and exports every
symbol include/mgx.h declares (no device calls)."""
import ctypes
import os
import re

import pytest

from mgmol_amd import LIB_PATH, mgx

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "mgx.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(mgx_[a-z0-9_]+)\s*\(", txt)))


def test_header_matches_binding_list():
    assert _declared() == sorted(mgx.EXPORTS)


def test_library_exports_every_symbol():
    if not os.path.exists(LIB_PATH):
        pytest.fail("libmgx.so not built: run python mgmol_amd/build.py")
    lib = ctypes.CDLL(LIB_PATH)
    missing = [s for s in _declared() if not hasattr(lib, s)]
    assert not missing, missing


def test_binding_signatures_load():
    L = mgx.lib()
    for s in mgx.EXPORTS:
        assert getattr(L, s) is not None


def test_gamma_host_function_matches_formula():
    import oracle as O
    h = (0.2, 0.2, 0.2)
    g = mgx.lib().mgx_gamma((ctypes.c_double * 3)(*h), 2, -0.05, -1.7)
    assert g == pytest.approx(O.gamma(list(h), 2, -0.05, -1.7), rel=1e-15)
