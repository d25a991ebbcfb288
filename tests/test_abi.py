"""CPU-side checks of the C-ABI library: it builds, loads, and exports every
entry point ``include/cadycore.h`` declares (no compute calls without a GPU)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "cadycore.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ca_[a-z_]+)\s*\(", text)))


@pytest.fixture(scope="module")
def lib():
    from cadycore import lib as L

    if not os.path.exists(L.LIB_PATH):
        from cadycore import build

        build.build()
    return ctypes.CDLL(L.LIB_PATH)


def test_header_declares_the_hook_set():
    syms = declared_symbols()
    for s in ("ca_create", "ca_remaining_tendency", "ca_implicit_tendency", "ca_update_jacobian", "ca_ldiv",
              "ca_correct_implicit_advection", "ca_set_precomputed", "ca_dss", "ca_step"):
        assert s in syms


def test_library_exports_every_declared_symbol(lib):
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_python_binding_covers_every_declared_symbol():
    from cadycore import lib as L

    assert sorted(L.EXPORTED) == declared_symbols()


def test_profile_count_is_callable_without_a_device(lib):
    lib.ca_profile_count.restype = ctypes.c_int32
    assert lib.ca_profile_count() >= 1


def test_create_rejects_bad_arguments_without_touching_the_device(lib):
    from cadycore import lib as L

    lib.ca_create.restype = ctypes.c_int
    g = L.CaGeometry()
    g.nlev = 1  # fewer than three levels is rejected before any HIP call
    g.ncol = 1
    p = L.CaParams()
    c = L.CaConfig()
    h = ctypes.c_void_p()
    assert lib.ca_create(ctypes.byref(g), ctypes.byref(p), ctypes.byref(c), ctypes.byref(h)) == 1
    assert lib.ca_create(None, None, None, ctypes.byref(h)) == 1
