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


def test_create_validates_tables_before_touching_the_device(lib):
    """Null tables, non-increasing faces and non-GLL quadrature tables are
    rejected by ca_create's argument checks, which run before any HIP call."""
    import numpy as np

    from cadycore import geometry as G
    from cadycore import lib as L
    from cadycore import params as P

    lib.ca_create.restype = ctypes.c_int
    lib.ca_last_error.restype = ctypes.c_char_p
    par = P.Params.from_toml()
    grid = G.cubed_sphere(1, G.uniform_faces(4, 30000.0), par.planet_radius, par.Omega)

    def create(edit):
        g, p, c, keep = L.create_args(grid, par, P.Numerics())
        edit(g, c, keep)
        h = ctypes.c_void_p()
        st = lib.ca_create(ctypes.byref(g), ctypes.byref(p), ctypes.byref(c), ctypes.byref(h))
        return st, lib.ca_last_error(None).decode(), h

    def null(field):
        return lambda g, c, keep: setattr(g, field, None)

    for f in ("z_face", "J_h", "gup", "gdn", "f3_num", "f12", "gll_w", "gll_D", "basis", "dss_ptr"):
        st, msg, h = create(null(f))
        assert st == 1 and h.value is None and "ca_create" in msg, (f, msg)

    def flat_faces(g, c, keep):
        z = keep["z"].copy()
        z[2] = z[1]
        keep["z2"] = z
        g.z_face = L._dp(z)

    st, msg, _ = create(flat_faces)
    assert st == 1 and "z_face" in msg

    def other_weights(g, c, keep):
        w = np.array([0.25, 0.75, 0.75, 0.25])
        keep["w2"] = w
        g.gll_w = L._dp(w)

    st, msg, _ = create(other_weights)
    assert st == 1 and "GLL" in msg

    def bad_upwinding(g, c, keep):
        c.upwinding = 2

    st, msg, _ = create(bad_upwinding)
    assert st == 1 and "upwinding" in msg
