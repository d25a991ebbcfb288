"""Python host side of the native dry-dycore step.

Binds ``libcadycore.so`` (C ABI in ``include/cadycore.h``) with ctypes and
mirrors the reference's tendency interface (``src/simulation/integrator.jl:
193-225``): ``remaining_tendency``, ``implicit_tendency``,
``update_jacobian`` (Wfact), ``ldiv``, ``correct_implicit_advection_tendency``
(T_post_imp!), ``set_implicit_precomputed_quantities`` (cache_imp!), ``dss``
and ``step`` (one ARS343 step of ``CTS.step!``).  States are float32 torch
tensors on the GPU in the native flat layout ``[rho | u1 | u2 | rho_e | u3]``.

There is no CPU fallback: constructing a :class:`DryDycore` without the
built library or without a GPU raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .geometry import Grid, gll
from .params import Numerics, Params

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcadycore.so")

CA_OK = 0
_STATUS = {0: "CA_OK", 1: "CA_ERR_ARG", 2: "CA_ERR_DEVICE", 3: "CA_ERR_NONFINITE"}


class CaGeometry(C.Structure):
    _fields_ = [
        ("nelem", C.c_int32), ("ncol", C.c_int32), ("nlev", C.c_int32), ("radius", C.c_double),
        ("z_face", C.POINTER(C.c_double)), ("J_h", C.POINTER(C.c_double)),
        ("gup", C.POINTER(C.c_double)), ("gdn", C.POINTER(C.c_double)),
        ("f3_num", C.POINTER(C.c_double)), ("f12", C.POINTER(C.c_double)),
        ("basis", C.POINTER(C.c_double)), ("n_dss", C.c_int32),
        ("dss_ptr", C.POINTER(C.c_int32)), ("dss_col", C.POINTER(C.c_int32)),
        ("dss_w", C.POINTER(C.c_double)), ("gll_w", C.POINTER(C.c_double)),
        ("gll_D", C.POINTER(C.c_double)),
    ]


class CaParams(C.Structure):
    _fields_ = [(n, C.c_double) for n in
                ("R_d", "cp_d", "cv_d", "T_0", "p_ref_theta", "grav", "T_surf_ref", "T_min_ref", "T_min_sgs")]


class CaConfig(C.Structure):
    _fields_ = [
        ("hyperdiff", C.c_int32), ("nu4_vort", C.c_double), ("nu4_scal", C.c_double),
        ("div_damp", C.c_double), ("upwinding", C.c_int32), ("rayleigh", C.c_int32),
        ("zd_rayleigh", C.c_double), ("alpha_uh", C.c_double), ("alpha_w", C.c_double),
        ("alpha_w_col", C.POINTER(C.c_double)), ("vert_diff", C.c_int32),
        ("D0_col", C.POINTER(C.c_double)), ("dt", C.c_double),
        ("sfc_flux", C.c_int32), ("Cd", C.c_double), ("Ch", C.c_double), ("gustiness", C.c_double),
        ("T_sfc_col", C.POINTER(C.c_double)),
        ("implicit_diffusion", C.c_int32), ("solve_iters", C.c_int32),
        ("vert_diff_model", C.c_int32), ("vert_diff_H", C.c_double), ("vert_diff_C_E", C.c_double),
    ]


class CaMember(C.Structure):
    _fields_ = [("nlev", C.c_int32), ("z_face", C.POINTER(C.c_double)), ("alpha_w", C.c_double),
                ("D0", C.c_double), ("f3", C.c_double)]


class CaLayout(C.Structure):
    _fields_ = [("offset", C.c_int64), ("s_v", C.c_int64), ("s_i", C.c_int64),
                ("s_j", C.c_int64), ("s_h", C.c_int64)]


class CaFieldVector(C.Structure):
    _fields_ = [("c", C.c_void_p), ("f", C.c_void_p), ("lay", CaLayout * 5)]


EXPORTED = (
    "ca_create", "ca_destroy", "ca_state_size", "ca_last_error", "ca_set_precomputed",
    "ca_remaining_tendency", "ca_implicit_tendency", "ca_update_jacobian", "ca_ldiv",
    "ca_correct_implicit_advection", "ca_dss", "ca_step", "ca_step_graph", "ca_pack_state", "ca_unpack_state",
    "ca_check_finite", "ca_profile_enable", "ca_profile_count", "ca_profile_get",
    "ca_set_tuning", "ca_get_tuning",
    "ca_ensemble_create", "ca_ensemble_destroy", "ca_ensemble_state_size", "ca_ensemble_member_offset",
    "ca_ensemble_last_error", "ca_ensemble_step", "ca_ensemble_step_graph", "ca_ensemble_set_tuning",
    "ca_ensemble_get_tuning", "ca_ensemble_autotune", "ca_ensemble_set_params", "ca_ensemble_adjoint",
    "ca_ensemble_adjoint_noalloc",
    "ca_dss_vjp", "ca_remaining_tendency_vjp", "ca_step_adjoint", "ca_step_record", "ca_step_pullback",
    "ca_set_column_params", "ca_set_debug_poison",
    "ca_fieldvector_native", "ca_fieldvector_vijfh", "ca_set_precomputed_quantities_fv",
    "ca_set_implicit_precomputed_quantities_fv", "ca_remaining_tendency_fv", "ca_implicit_tendency_fv",
    "ca_update_jacobian_fv", "ca_ldiv_fv", "ca_correct_implicit_advection_fv", "ca_dss_fv",
)
CA_TUNE_NEWTON, CA_TUNE_EXPLICIT_CHUNK, CA_TUNE_DSS_CHUNK, CA_TUNE_DSS_KIND, CA_TUNE_ADJ_CHUNK, CA_TUNE_LAP_CHUNK = 0, 1, 2, 3, 4, 5
CA_ENS_LANES, CA_ENS_ORDER = 0, 1
_NOT_STATUS = ("ca_state_size", "ca_last_error", "ca_profile_count", "ca_ensemble_state_size",
               "ca_ensemble_member_offset", "ca_ensemble_last_error")


def load_library(path: str = LIB_PATH):
    if not os.path.exists(path):
        raise RuntimeError(f"libcadycore.so not built ({path}); run __graft_entry__.build()")
    lib = C.CDLL(path)
    vp = C.c_void_p
    lib.ca_create.argtypes = [C.POINTER(CaGeometry), C.POINTER(CaParams), C.POINTER(CaConfig), C.POINTER(vp)]
    lib.ca_destroy.argtypes = [vp]
    lib.ca_state_size.argtypes = [vp]
    lib.ca_state_size.restype = C.c_int64
    lib.ca_last_error.argtypes = [vp]
    lib.ca_last_error.restype = C.c_char_p
    for nm in ("ca_set_precomputed", "ca_dss"):
        getattr(lib, nm).argtypes = [vp, vp, vp]
    for nm in ("ca_remaining_tendency", "ca_implicit_tendency", "ca_ldiv", "ca_correct_implicit_advection"):
        getattr(lib, nm).argtypes = [vp, vp, vp, vp]
    lib.ca_update_jacobian.argtypes = [vp, vp, C.c_double, vp]
    lib.ca_step.argtypes = [vp, vp, C.c_double, C.c_int32, vp]
    lib.ca_step_graph.argtypes = [vp, vp, C.c_double, C.c_int32, vp]
    fvp = C.POINTER(CaFieldVector)
    lib.ca_fieldvector_native.argtypes = [vp, vp, fvp]
    lib.ca_fieldvector_vijfh.argtypes = [vp, vp, vp, fvp]
    lib.ca_set_precomputed_quantities_fv.argtypes = [vp, fvp, C.c_double, vp]
    lib.ca_set_implicit_precomputed_quantities_fv.argtypes = [vp, fvp, C.c_double, vp]
    lib.ca_remaining_tendency_fv.argtypes = [vp, fvp, fvp, fvp, C.c_double, vp]
    lib.ca_implicit_tendency_fv.argtypes = [vp, fvp, fvp, C.c_double, vp]
    lib.ca_update_jacobian_fv.argtypes = [vp, fvp, C.c_double, C.c_double, vp]
    lib.ca_ldiv_fv.argtypes = [vp, fvp, fvp, vp]
    lib.ca_correct_implicit_advection_fv.argtypes = [vp, fvp, fvp, C.c_double, vp]
    lib.ca_dss_fv.argtypes = [vp, fvp, C.c_double, vp]
    lib.ca_pack_state.argtypes = [vp, vp, C.POINTER(CaLayout), vp, vp]
    lib.ca_unpack_state.argtypes = [vp, vp, C.POINTER(CaLayout), vp, vp]
    lib.ca_check_finite.argtypes = [vp, vp, C.POINTER(C.c_int32), vp]
    lib.ca_profile_enable.argtypes = [vp, C.c_int32]
    lib.ca_set_debug_poison.argtypes = [vp, C.c_int32]
    lib.ca_profile_count.argtypes = []
    lib.ca_set_tuning.argtypes = [vp, C.c_int32, C.c_int32]
    lib.ca_get_tuning.argtypes = [vp, C.c_int32, C.POINTER(C.c_int32)]
    lib.ca_profile_get.argtypes = [vp, C.c_int32, C.POINTER(C.c_char_p), C.POINTER(C.c_int64),
                                   C.POINTER(C.c_double)]
    lib.ca_ensemble_create.argtypes = [C.POINTER(CaParams), C.POINTER(CaConfig), C.c_int32, C.POINTER(CaMember),
                                       C.POINTER(vp)]
    lib.ca_ensemble_destroy.argtypes = [vp]
    lib.ca_ensemble_state_size.argtypes = [vp]
    lib.ca_ensemble_state_size.restype = C.c_int64
    lib.ca_ensemble_member_offset.argtypes = [vp, C.c_int32]
    lib.ca_ensemble_member_offset.restype = C.c_int64
    lib.ca_ensemble_last_error.argtypes = [vp]
    lib.ca_ensemble_last_error.restype = C.c_char_p
    lib.ca_ensemble_step.argtypes = [vp, vp, C.c_double, C.c_int32, vp]
    lib.ca_ensemble_step_graph.argtypes = [vp, vp, C.c_double, C.c_int32, vp]
    lib.ca_ensemble_set_tuning.argtypes = [vp, C.c_int32, C.c_int32]
    lib.ca_ensemble_get_tuning.argtypes = [vp, C.c_int32, C.POINTER(C.c_int32)]
    lib.ca_ensemble_autotune.argtypes = [vp, vp, C.c_double, C.c_int32, vp, C.POINTER(C.c_double)]
    dpp = C.POINTER(C.c_double)
    lib.ca_ensemble_set_params.argtypes = [vp, dpp, dpp]
    lib.ca_ensemble_adjoint.argtypes = [vp, vp, vp, C.c_double, C.c_int32, C.c_int32, vp, vp, vp, vp]
    lib.ca_ensemble_adjoint_noalloc.argtypes = [vp, vp, vp, C.c_double, C.c_int32, C.c_int32, vp, vp, vp, vp]
    lib.ca_dss_vjp.argtypes = [vp, vp, vp]
    lib.ca_remaining_tendency_vjp.argtypes = [vp, vp, vp, vp, vp, vp]
    lib.ca_step_adjoint.argtypes = [vp, vp, vp, C.c_double, C.c_int32, C.c_int32, vp, vp, vp]
    lib.ca_step_record.argtypes = [vp, vp, C.c_double, C.c_int32, C.c_int32, vp]
    lib.ca_step_pullback.argtypes = [vp, vp, vp, vp, vp]
    lib.ca_set_column_params.argtypes = [vp, dpp, dpp]
    for nm in EXPORTED:
        if nm not in _NOT_STATUS:
            getattr(lib, nm).restype = C.c_int
    lib.ca_profile_count.restype = C.c_int32
    return lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def create_args(grid: Grid, params: Params, numerics: Numerics, alpha_w_col=None, D0_col=None, T_sfc_col=None):
    """The ca_geometry / ca_params / ca_config structs of ca_create for a grid,
    and the arrays they point into (keep them alive while the structs are used)."""
    keep = {}

    def arr(name, a, dt=np.float64):
        a = np.ascontiguousarray(a, dtype=dt)
        keep[name] = a
        return a

    _, wq, D = gll()
    g = CaGeometry()
    g.nelem, g.ncol, g.nlev, g.radius = grid.nelem, grid.ncol, grid.nlev, grid.radius
    g.z_face = _dp(arr("z", grid.z_face))
    g.J_h = _dp(arr("J", grid.J_h))
    g.gup = _dp(arr("gup", grid.gup))
    g.gdn = _dp(arr("gdn", grid.gdn))
    g.f3_num = _dp(arr("f3", grid.f3_num))
    g.f12 = _dp(arr("f12", grid.f12))
    g.basis = _dp(arr("b", grid.basis))
    g.n_dss = len(grid.dss_ptr) - 1
    g.dss_ptr = arr("dp", grid.dss_ptr, np.int32).ctypes.data_as(C.POINTER(C.c_int32))
    g.dss_col = arr("dc", grid.dss_col, np.int32).ctypes.data_as(C.POINTER(C.c_int32))
    g.dss_w = _dp(arr("dw", grid.dss_w))
    g.gll_w = _dp(arr("w", wq))
    g.gll_D = _dp(arr("D", D))
    p = CaParams(params.R_d, params.cp_d, params.cv_d, params.T_0, params.p_ref_theta, params.grav,
                 params.T_surf_ref, params.T_min_ref, params.T_min_sgs)
    c = CaConfig()
    c.hyperdiff = int(bool(numerics.hyperdiff) and grid.nelem > 0)
    nu_v, nu_s = numerics.nu4(grid.h) if grid.nelem > 0 else (0.0, 0.0)
    c.nu4_vort, c.nu4_scal, c.div_damp = nu_v, nu_s, numerics.div_damping
    c.upwinding = int(numerics.upwinding)
    c.rayleigh = int(bool(numerics.rayleigh))
    c.zd_rayleigh, c.alpha_uh, c.alpha_w = params.zd_rayleigh, params.alpha_rayleigh_uh, params.alpha_rayleigh_w
    if alpha_w_col is not None:
        c.alpha_w_col = _dp(arr("aw", alpha_w_col))
    c.vert_diff = int(bool(numerics.vert_diff))
    c.vert_diff_model = numerics.vert_diff_kind()
    c.vert_diff_H = float(numerics.vert_diff_H)
    c.vert_diff_C_E = float(numerics.vert_diff_C_E)
    if numerics.vert_diff and c.vert_diff_model == 0:
        if D0_col is None:
            raise ValueError("vertical diffusion needs D0_col")
        c.D0_col = _dp(arr("D0", D0_col))
    c.dt = numerics.dt
    c.sfc_flux = int(bool(numerics.sfc_flux))
    if numerics.sfc_flux:
        if T_sfc_col is None:
            raise ValueError("surface fluxes need T_sfc_col")
        c.Cd, c.Ch, c.gustiness = numerics.Cd, numerics.Ch, numerics.gustiness
        c.T_sfc_col = _dp(arr("Tsfc", T_sfc_col))
    c.implicit_diffusion = int(bool(numerics.implicit_diffusion))
    c.solve_iters = int(numerics.solve_iters)
    return g, p, c, keep


class DryDycore:
    """One native dry-dycore instance on the current GPU.

    ``grid``/``params``/``numerics`` describe the configuration;
    ``alpha_w_col`` and ``D0_col`` give per-column Rayleigh sponge strengths
    and vertical diffusivities (ensembles and calibration); ``T_sfc_col`` the
    per-column surface temperature of the surface fluxes.
    """

    def __init__(self, grid: Grid, params: Params, numerics: Numerics, alpha_w_col=None, D0_col=None,
                 device=None, lib_path: str = LIB_PATH, T_sfc_col=None):
        import torch

        if not torch.cuda.is_available():
            raise RuntimeError("DryDycore needs a GPU (no CPU fallback)")
        self.torch = torch
        self.device = torch.device(device or "cuda:0")
        torch.cuda.set_device(self.device)
        self.lib = load_library(lib_path)
        self.grid, self.params, self.numerics = grid, params, numerics
        g, p, c, keep = create_args(grid, params, numerics, alpha_w_col, D0_col, T_sfc_col)
        h = C.c_void_p()
        st = self.lib.ca_create(C.byref(g), C.byref(p), C.byref(c), C.byref(h))
        if st != CA_OK:
            msg = self.lib.ca_last_error(None)
            raise RuntimeError(f"ca_create failed: {_STATUS.get(st, st)}: {msg.decode() if msg else ''}")
        self.h = h
        self.S = int(self.lib.ca_state_size(h))
        self._keep = keep

    # ------------------------------------------------------------ plumbing
    def close(self):
        if getattr(self, "h", None):
            self.lib.ca_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _stream(self):
        return C.c_void_p(self.torch.cuda.current_stream(self.device).cuda_stream)

    def _ptr(self, t):
        assert t.is_cuda and t.dtype == self.torch.float32 and t.is_contiguous() and t.numel() == self.S
        return C.c_void_p(t.data_ptr())

    def _check(self, st, what):
        if st != CA_OK:
            msg = self.lib.ca_last_error(self.h)
            raise RuntimeError(f"{what}: {_STATUS.get(st, st)}: {msg.decode() if msg else ''}")

    def state(self, Y_np):
        """Upload a numpy state (float64 or float32) as a device tensor."""
        return self.torch.tensor(np.asarray(Y_np, dtype=np.float32), device=self.device)

    def zeros(self):
        return self.torch.zeros(self.S, dtype=self.torch.float32, device=self.device)

    # ------------------------------------------------- reference interface
    def set_implicit_precomputed_quantities(self, Y):
        self._check(self.lib.ca_set_precomputed(self.h, self._ptr(Y), self._stream()), "cache_imp!")

    def remaining_tendency(self, Yt, Y):
        self._check(self.lib.ca_remaining_tendency(self.h, self._ptr(Yt), self._ptr(Y), self._stream()),
                    "remaining_tendency!")
        return Yt

    def implicit_tendency(self, Yt, Y):
        self._check(self.lib.ca_implicit_tendency(self.h, self._ptr(Yt), self._ptr(Y), self._stream()),
                    "implicit_tendency!")

    def update_jacobian(self, Y, dtgamma):
        self._check(self.lib.ca_update_jacobian(self.h, self._ptr(Y), float(dtgamma), self._stream()),
                    "update_jacobian!")

    def ldiv(self, dY, R):
        self._check(self.lib.ca_ldiv(self.h, self._ptr(dY), self._ptr(R), self._stream()), "ldiv!")

    def correct_implicit_advection_tendency(self, Yt, Y):
        self._check(self.lib.ca_correct_implicit_advection(self.h, self._ptr(Yt), self._ptr(Y), self._stream()),
                    "correct_implicit_advection_tendency!")

    def dss(self, Y):
        self._check(self.lib.ca_dss(self.h, self._ptr(Y), self._stream()), "dss!")

    def step(self, Y, dt, nsteps=1):
        self._check(self.lib.ca_step(self.h, self._ptr(Y), float(dt), int(nsteps), self._stream()), "step!")

    def step_graph(self, Y, dt, nsteps, stream):
        """``step`` replayed from a captured hipGraph on ``stream`` (a
        torch.cuda.Stream other than the default one)."""
        self._check(self.lib.ca_step_graph(self.h, self._ptr(Y), float(dt), int(nsteps),
                                           C.c_void_p(stream.cuda_stream)), "step (graph)")

    # ------------------------------------- hooks on ClimaCore FieldVectors
    def fieldvector(self, Y=None, c=None, f=None):
        """A ca_fieldvector of a native state tensor ``Y``, or of VIJFH parent
        arrays ``c`` (Julia (Nv, 4, 4, 4, Nh)) and ``f`` ((Nv + 1, 4, 4, 1, Nh))."""
        fv = CaFieldVector()
        if Y is not None:
            self._check(self.lib.ca_fieldvector_native(self.h, self._ptr(Y), C.byref(fv)), "fieldvector")
            fv._keep = (Y,)
        else:
            for t in (c, f):
                assert t.is_cuda and t.dtype == self.torch.float32 and t.is_contiguous()
            self._check(self.lib.ca_fieldvector_vijfh(self.h, C.c_void_p(c.data_ptr()), C.c_void_p(f.data_ptr()),
                                                      C.byref(fv)), "fieldvector")
            fv._keep = (c, f)
        return fv

    def to_vijfh(self, Y):
        """Native state tensor -> (c, f) VIJFH parent arrays (C-order shapes
        (Nh, 4, 4, 4, Nv) and (Nh, 1, 4, 4, Nv + 1): Julia's column-major
        (Nv, Ni, Nj, F, Nh))."""
        g = self.grid
        nl, nh = g.nlev, g.ncol // 16
        nC = nl * g.ncol
        cen = Y[:4 * nC].reshape(4, nl, nh, 4, 4)           # (f, v, h, j, i)
        c = cen.permute(2, 0, 3, 4, 1).contiguous()          # (h, f, j, i, v)
        fac = Y[4 * nC:].reshape(1, nl + 1, nh, 4, 4)
        f = fac.permute(2, 0, 3, 4, 1).contiguous()
        return c, f

    def from_vijfh(self, c, f):
        g = self.grid
        nl = g.nlev
        cen = c.permute(1, 4, 0, 2, 3).reshape(-1)           # (f, v, h, j, i)
        fac = f.permute(1, 4, 0, 2, 3).reshape(-1)
        return self.torch.cat([cen, fac]).contiguous()

    def set_precomputed_quantities_fv(self, Y, t=0.0):
        self._check(self.lib.ca_set_precomputed_quantities_fv(self.h, C.byref(Y), float(t), self._stream()), "cache!")

    def set_implicit_precomputed_quantities_fv(self, Y, t=0.0):
        self._check(self.lib.ca_set_implicit_precomputed_quantities_fv(self.h, C.byref(Y), float(t), self._stream()),
                    "cache_imp!")

    def remaining_tendency_fv(self, Yt, Yt_lim, Y, t=0.0):
        lim = C.byref(Yt_lim) if Yt_lim is not None else None
        self._check(self.lib.ca_remaining_tendency_fv(self.h, C.byref(Yt), lim, C.byref(Y), float(t), self._stream()),
                    "remaining_tendency!")

    def implicit_tendency_fv(self, Yt, Y, t=0.0):
        self._check(self.lib.ca_implicit_tendency_fv(self.h, C.byref(Yt), C.byref(Y), float(t), self._stream()),
                    "implicit_tendency!")

    def update_jacobian_fv(self, Y, dtgamma, t=0.0):
        self._check(self.lib.ca_update_jacobian_fv(self.h, C.byref(Y), float(dtgamma), float(t), self._stream()),
                    "update_jacobian!")

    def ldiv_fv(self, dY, R):
        self._check(self.lib.ca_ldiv_fv(self.h, C.byref(dY), C.byref(R), self._stream()), "ldiv!")

    def correct_implicit_advection_tendency_fv(self, Yt, Y, t=0.0):
        self._check(self.lib.ca_correct_implicit_advection_fv(self.h, C.byref(Yt), C.byref(Y), float(t),
                                                              self._stream()), "correct_implicit_advection_tendency!")

    def dss_fv(self, Y, t=0.0):
        self._check(self.lib.ca_dss_fv(self.h, C.byref(Y), float(t), self._stream()), "dss!")

    # ------------------------------------------------------- reverse mode
    def dss_vjp(self, Yb):
        """Yb <- DSS^T Yb (in place): the transpose of dss!."""
        self._check(self.lib.ca_dss_vjp(self.h, self._ptr(Yb), self._stream()), "dss! reverse")

    def remaining_tendency_vjp(self, Yb, Ytb, Y, D0b=None):
        """Yb <- (d T_exp / d Y)^T Ytb at Y, and (vertical diffusion) D0b <-
        d <Ytb, T_exp> / d D0 per column; sphere grids."""
        d0 = C.c_void_p(D0b.data_ptr()) if D0b is not None else None
        self._check(self.lib.ca_remaining_tendency_vjp(self.h, self._ptr(Yb), self._ptr(Ytb), self._ptr(Y), d0,
                                                       self._stream()), "remaining_tendency! reverse")

    def step_adjoint(self, Y0, YbN, dt, nsteps, every=8, D0b=None):
        """Reverse mode of ``step(Y0, dt, nsteps)``: given the cotangent YbN of
        the state after the steps, returns the cotangent of Y0 (and fills
        D0b, [ncol] float32, with the D0 cotangent); sphere grids."""
        Yb0 = self.zeros()
        d0 = C.c_void_p(D0b.data_ptr()) if D0b is not None else None
        self._check(self.lib.ca_step_adjoint(self.h, self._ptr(Y0), self._ptr(YbN), float(dt), int(nsteps),
                                             int(every), self._ptr(Yb0), d0, self._stream()), "step adjoint")
        return Yb0

    def step_record(self, Y, dt, nsteps, every=8):
        """``step`` that keeps its checkpoints for ``step_pullback``."""
        self._check(self.lib.ca_step_record(self.h, self._ptr(Y), float(dt), int(nsteps), int(every), self._stream()),
                    "step record")

    def step_pullback(self, YbN, D0b=None):
        """Pullback of the last ``step_record``: returns the cotangent of its
        input (and fills D0b, [ncol] float32)."""
        Yb0 = self.zeros()
        d0 = C.c_void_p(D0b.data_ptr()) if D0b is not None else None
        self._check(self.lib.ca_step_pullback(self.h, self._ptr(YbN), self._ptr(Yb0), d0, self._stream()),
                    "step pullback")
        return Yb0

    def set_column_params(self, alpha_w_col=None, D0_col=None):
        def arr(x):
            if x is None:
                return None
            a = np.ascontiguousarray(np.asarray(x, np.float64).reshape(-1))
            assert a.size == self.grid.ncol
            self._keep[f"p{id(a)}"] = a
            return _dp(a)
        self._check(self.lib.ca_set_column_params(self.h, arr(alpha_w_col), arr(D0_col)), "set_column_params")

    def check_finite(self, Y):
        flag = C.c_int32(0)
        st = self.lib.ca_check_finite(self.h, self._ptr(Y), C.byref(flag), self._stream())
        return st == CA_OK and flag.value == 0

    def set_debug_poison(self, on=True):
        """NaN-fill the handle's work buffers at every call (fill_with_nans!)."""
        self._check(self.lib.ca_set_debug_poison(self.h, int(on)), "set_debug_poison")

    def set_tuning(self, knob, value):
        self._check(self.lib.ca_set_tuning(self.h, int(knob), int(value)), "set_tuning")

    def get_tuning(self, knob):
        v = C.c_int32()
        self._check(self.lib.ca_get_tuning(self.h, int(knob), C.byref(v)), "get_tuning")
        return v.value

    # ------------------------------------------------------------ profiling
    def profile_enable(self, on=True):
        self._check(self.lib.ca_profile_enable(self.h, int(on)), "profile_enable")

    def profile(self):
        out = {}
        for i in range(self.lib.ca_profile_count()):
            name = C.c_char_p()
            n = C.c_int64()
            ms = C.c_double()
            self._check(self.lib.ca_profile_get(self.h, i, C.byref(name), C.byref(n), C.byref(ms)), "profile_get")
            out[name.value.decode()] = (int(n.value), float(ms.value))
        return out
