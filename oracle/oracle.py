"""ctypes binding of the CPU restatement (``ca_oracle.c``).

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py`` -- never by the product package.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_BUILD = os.path.join(_HERE, "_build")


def build(force: bool = False):
    if force or not os.path.exists(os.path.join(_BUILD, "liboracle.so")) or not os.path.exists(
        os.path.join(_BUILD, "liboracle_f32.so")
    ):
        subprocess.check_call(["make", "-s", "-C", _HERE], stdout=subprocess.DEVNULL)


class _Geom(C.Structure):
    _fields_ = [
        ("nelem", C.c_int32), ("ncol", C.c_int32), ("nlev", C.c_int32),
        ("radius", C.c_double),
        ("z_face", C.POINTER(C.c_double)), ("J_h", C.POINTER(C.c_double)),
        ("gup", C.POINTER(C.c_double)), ("gdn", C.POINTER(C.c_double)),
        ("f3_num", C.POINTER(C.c_double)), ("f12", C.POINTER(C.c_double)),
        ("basis", C.POINTER(C.c_double)),
        ("n_dss", C.c_int32), ("dss_ptr", C.POINTER(C.c_int32)),
        ("dss_col", C.POINTER(C.c_int32)), ("dss_w", C.POINTER(C.c_double)),
        ("gll_w", C.POINTER(C.c_double)), ("gll_D", C.POINTER(C.c_double)),
    ]


class _Num(C.Structure):
    _fields_ = [
        ("R_d", C.c_double), ("cp_d", C.c_double), ("cv_d", C.c_double), ("T_0", C.c_double),
        ("p_ref_theta", C.c_double), ("grav", C.c_double), ("T_surf_ref", C.c_double),
        ("T_min_ref", C.c_double), ("T_min_sgs", C.c_double),
        ("hyperdiff", C.c_int32), ("nu4_vort", C.c_double), ("nu4_scal", C.c_double),
        ("div_damp", C.c_double), ("upwinding", C.c_int32), ("rayleigh", C.c_int32),
        ("zd_rayleigh", C.c_double), ("alpha_uh", C.c_double), ("alpha_w", C.c_double),
        ("alpha_w_col", C.POINTER(C.c_double)), ("vert_diff", C.c_int32),
        ("D0_col", C.POINTER(C.c_double)), ("dt", C.c_double),
        ("sfc_flux", C.c_int32), ("Cd", C.c_double), ("Ch", C.c_double), ("gustiness", C.c_double),
        ("T_sfc_col", C.POINTER(C.c_double)),
        ("implicit_diffusion", C.c_int32), ("solve_iters", C.c_int32),
        ("vdiff_kind", C.c_int32), ("vdiff_H", C.c_double), ("vdiff_C_E", C.c_double),
    ]


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


class Oracle:
    """One oracle context: grid + constants + numerics.

    ``f32=True`` loads the Float32/OpenMP build (CPU timing baseline only).
    ``lib_path`` loads another build of the same source (the pinning tests'
    mutation run plants a mistake in a copy of it to show the checks catch it).
    """

    def __init__(self, grid, params, numerics, alpha_w_col=None, D0_col=None, f32=False,
                 lib_path=None, T_sfc_col=None):
        if lib_path is None:
            build()
            lib_path = os.path.join(_BUILD, "liboracle_f32.so" if f32 else "liboracle.so")
        lib = C.CDLL(lib_path)
        self.lib = lib
        self.f32 = f32
        self.dtype = np.float32 if f32 else np.float64
        rp = C.POINTER(C.c_float if f32 else C.c_double)
        self._rp = rp
        lib.ora_create.argtypes = [C.POINTER(_Geom), C.POINTER(_Num), C.POINTER(C.c_void_p)]
        lib.ora_destroy.argtypes = [C.c_void_p]
        lib.ora_state_size.argtypes = [C.c_void_p]
        lib.ora_state_size.restype = C.c_int64
        for name in ("ora_set_precomputed", "ora_dss"):
            getattr(lib, name).argtypes = [C.c_void_p, rp]
        for name in ("ora_remaining_tendency", "ora_implicit_tendency", "ora_correct_implicit"):
            getattr(lib, name).argtypes = [C.c_void_p, rp, rp]
        lib.ora_update_jacobian.argtypes = [C.c_void_p, rp, C.c_double]
        lib.ora_ldiv.argtypes = [C.c_void_p, rp, rp]
        lib.ora_step.argtypes = [C.c_void_p, rp, C.c_double, C.c_int32]
        lib.ora_set_storage_rounding.argtypes = [C.c_void_p, C.c_int32]
        d = C.POINTER(C.c_double)
        lib.ora_get_precomputed.argtypes = [C.c_void_p, d, d, d, d]
        lib.ora_get_laplacians.argtypes = [C.c_void_p, d]
        for name in ("ora_op_grad", "ora_op_wgrad", "ora_op_curl12", "ora_op_wcurl12"):
            getattr(lib, name).argtypes = [C.c_void_p, C.c_int, d, d, d]
        for name in ("ora_op_div", "ora_op_wdiv", "ora_op_curl3", "ora_op_wcurl3"):
            getattr(lib, name).argtypes = [C.c_void_p, C.c_int, d, d, d]
        lib.ora_column_blocks.argtypes = [C.c_void_p, C.c_int] + [d] * 7
        lib.ora_column_diffusion_blocks.argtypes = [C.c_void_p, C.c_int, d, d]
        lib.ora_column_uh_blocks.argtypes = [C.c_void_p, C.c_int, d]

        from_geom = _import_geometry()
        xq, wq, D = from_geom.gll()
        g = grid
        self.grid = g
        keep = {}

        def arr(name, a, dt=np.float64):
            a = np.ascontiguousarray(a, dtype=dt)
            keep[name] = a
            return a

        geom = _Geom()
        geom.nelem = g.nelem
        geom.ncol = g.ncol
        geom.nlev = g.nlev
        geom.radius = g.radius
        geom.z_face = _dp(arr("z", g.z_face))
        geom.J_h = _dp(arr("J", g.J_h))
        geom.gup = _dp(arr("gup", g.gup))
        geom.gdn = _dp(arr("gdn", g.gdn))
        geom.f3_num = _dp(arr("f3", g.f3_num))
        geom.f12 = _dp(arr("f12", g.f12))
        geom.basis = _dp(arr("basis", g.basis))
        geom.n_dss = len(g.dss_ptr) - 1
        geom.dss_ptr = _ip(arr("dp", g.dss_ptr, np.int32))
        geom.dss_col = _ip(arr("dc", g.dss_col, np.int32))
        geom.dss_w = _dp(arr("dw", g.dss_w))
        geom.gll_w = _dp(arr("w", wq))
        geom.gll_D = _dp(arr("D", D))
        p, n = params, numerics
        num = _Num()
        num.R_d, num.cp_d, num.cv_d, num.T_0 = p.R_d, p.cp_d, p.cv_d, p.T_0
        num.p_ref_theta, num.grav = p.p_ref_theta, p.grav
        num.T_surf_ref, num.T_min_ref, num.T_min_sgs = p.T_surf_ref, p.T_min_ref, p.T_min_sgs
        num.hyperdiff = int(bool(n.hyperdiff) and g.nelem > 0)
        nu_v, nu_s = n.nu4(g.h) if g.nelem > 0 else (0.0, 0.0)
        num.nu4_vort, num.nu4_scal, num.div_damp = nu_v, nu_s, n.div_damping
        num.upwinding = int(n.upwinding)
        num.rayleigh = int(bool(n.rayleigh))
        num.zd_rayleigh, num.alpha_uh, num.alpha_w = p.zd_rayleigh, p.alpha_rayleigh_uh, p.alpha_rayleigh_w
        if alpha_w_col is not None:
            num.alpha_w_col = _dp(arr("aw", alpha_w_col))
        num.vert_diff = int(bool(n.vert_diff))
        num.vdiff_kind = n.vert_diff_kind() if hasattr(n, "vert_diff_kind") else 0
        num.vdiff_H = float(getattr(n, "vert_diff_H", float("inf")))
        num.vdiff_C_E = float(getattr(n, "vert_diff_C_E", 0.044))
        if n.vert_diff and num.vdiff_kind == 0:
            if D0_col is None:
                raise ValueError("vertical diffusion needs D0_col")
            num.D0_col = _dp(arr("D0", D0_col))
        num.dt = n.dt
        num.sfc_flux = int(bool(getattr(n, "sfc_flux", False)))
        if num.sfc_flux:
            if T_sfc_col is None:
                raise ValueError("surface fluxes need T_sfc_col")
            num.Cd, num.Ch, num.gustiness = n.Cd, n.Ch, n.gustiness
            num.T_sfc_col = _dp(arr("Tsfc", T_sfc_col))
        num.implicit_diffusion = int(bool(getattr(n, "implicit_diffusion", False)))
        num.solve_iters = int(getattr(n, "solve_iters", 1))
        self._keep = keep
        self._geom = geom
        self._num = num
        h = C.c_void_p()
        if lib.ora_create(C.byref(geom), C.byref(num), C.byref(h)) != 0:
            raise RuntimeError("ora_create failed")
        self.h = h
        self.S = lib.ora_state_size(h)

    def __del__(self):
        try:
            self.lib.ora_destroy(self.h)
        except Exception:
            pass

    def _v(self, a):
        assert a.dtype == self.dtype and a.flags.c_contiguous
        return a.ctypes.data_as(self._rp)

    def new(self):
        return np.zeros(self.S, dtype=self.dtype)

    def set_precomputed(self, Y):
        self.lib.ora_set_precomputed(self.h, self._v(Y))

    def remaining_tendency(self, Y):
        Yt = self.new()
        self.lib.ora_remaining_tendency(self.h, self._v(Yt), self._v(Y))
        return Yt

    def implicit_tendency(self, Y):
        Yt = self.new()
        self.lib.ora_implicit_tendency(self.h, self._v(Yt), self._v(Y))
        return Yt

    def update_jacobian(self, Y, dtgamma):
        self.lib.ora_update_jacobian(self.h, self._v(Y), float(dtgamma))

    def ldiv(self, R):
        dY = self.new()
        R = np.ascontiguousarray(R, dtype=self.dtype)
        self.lib.ora_ldiv(self.h, self._v(dY), self._v(R))
        return dY

    def correct_implicit(self, Y):
        Yt = self.new()
        self.lib.ora_correct_implicit(self.h, self._v(Yt), self._v(Y))
        return Yt

    def dss(self, Y):
        self.lib.ora_dss(self.h, self._v(Y))

    def step(self, Y, dt, nsteps=1):
        self.lib.ora_step(self.h, self._v(Y), float(dt), int(nsteps))

    def set_storage_rounding(self, on=True):
        """Round stage states, tendencies and the new state to float32 inside
        step(): the error of float32 storage alone (float64 build)."""
        self.lib.ora_set_storage_rounding(self.h, 1 if on else 0)

    def precomputed(self):
        nC = self.grid.nlev * self.grid.ncol
        out = [np.zeros(nC) for _ in range(4)]
        self.lib.ora_get_precomputed(self.h, *[_dp(o) for o in out])
        return out

    def laplacians(self):
        nC = self.grid.nlev * self.grid.ncol
        out = np.zeros(4 * nC)
        self.lib.ora_get_laplacians(self.h, _dp(out))
        return out.reshape(4, self.grid.nlev, self.grid.ncol)

    def op(self, name, k, *args):
        nc = self.grid.ncol
        ins = [np.ascontiguousarray(a, dtype=np.float64) for a in args]
        nout = 2 if name in ("grad", "wgrad", "curl12", "wcurl12") else 1
        outs = [np.zeros(nc) for _ in range(nout)]
        getattr(self.lib, "ora_op_" + name)(self.h, k, *[_dp(a) for a in ins], *[_dp(o) for o in outs])
        return outs if nout == 2 else outs[0]

    def column_blocks(self, col):
        nl = self.grid.nlev
        Aru, Aeu = np.zeros(2 * nl), np.zeros(2 * nl)
        Aur, Aue = np.zeros(2 * (nl + 1)), np.zeros(2 * (nl + 1))
        lo, di, up = np.zeros(nl + 1), np.zeros(nl + 1), np.zeros(nl + 1)
        self.lib.ora_column_blocks(self.h, col, *[_dp(a) for a in (Aru, Aeu, Aur, Aue, lo, di, up)])
        Aee, Ahh = np.zeros(3 * nl), np.zeros(3 * nl)
        self.lib.ora_column_diffusion_blocks(self.h, col, _dp(Aee), _dp(Ahh))
        Auh = np.zeros(4 * (nl + 1))
        self.lib.ora_column_uh_blocks(self.h, col, _dp(Auh))
        return dict(Aru=Aru.reshape(nl, 2), Aeu=Aeu.reshape(nl, 2), Aur=Aur.reshape(nl + 1, 2),
                    Aue=Aue.reshape(nl + 1, 2), lo=lo, di=di, up=up, Aee=Aee.reshape(nl, 3),
                    Ahh=Ahh.reshape(nl, 3), Auh=Auh.reshape(nl + 1, 4))


def _import_geometry():
    import importlib
    import sys

    pkg_root = os.path.join(os.path.dirname(_HERE), "climaatmos.jl_amd")
    if pkg_root not in sys.path:
        sys.path.insert(0, pkg_root)
    return importlib.import_module("cadycore.geometry")


def tableau():
    """The ARS343 tableau the oracle's step uses: (aE, aI, b, c)."""
    build()
    lib = C.CDLL(os.path.join(_BUILD, "liboracle.so"))
    aE, aI, b, c = np.zeros(16), np.zeros(16), np.zeros(4), np.zeros(4)
    lib.ora_tableau(_dp(aE), _dp(aI), _dp(b), _dp(c))
    return aE.reshape(4, 4), aI.reshape(4, 4), b, c
