"""Batched single-column ensembles of uneven vertical sizes (``ca_ensemble_*``).

Each member is an independent column (the reference's ``ColumnGrid``,
``src/simulation/grids.jl:110-129``) with its own face heights, Rayleigh
coefficient ``alpha_w`` and vertical diffusivity ``D0``, stepped with the
same ARS343 step as :meth:`cadycore.lib.DryDycore.step` on a column grid.
All members advance together in one persistent launch, which can be
captured in a hipGraph and replayed.  The reference has no counterpart: it
runs one column per simulation.

The state is one flat float32 tensor; member ``m`` occupies
``[off[m], off[m] + 5 nlev_m + 1)`` as ``[rho | u1 | u2 | rho_e | u3]``.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import geometry as G
from . import initial as I
from .lib import (CA_ENS_LANES, CA_ENS_ORDER, CA_OK, LIB_PATH, _STATUS, CaConfig, CaMember, CaParams,
                  load_library)
from .params import Numerics, Params


@dataclass
class Member:
    z_face: np.ndarray
    alpha_w: float = 0.0
    D0: float = 0.0
    f3: float = 0.0

    @property
    def nlev(self) -> int:
        return len(self.z_face) - 1


def member_faces(z_elem: int, z_max: float = 30000.0, dz_bottom: float = 500.0, dz_fallback: float = 30.0):
    """Stretched faces with the column default dz_bottom = 500 m when it is
    thinner than the mean layer, else the z63 production value of 30 m."""
    dz = dz_bottom if dz_bottom < z_max / z_elem else dz_fallback
    return G.tanh_stretched_faces(z_elem, z_max, dz)


def c3_members(params: Params, n: int = 512, seed: int = 3):
    """The ragged ensemble workload: z_elem ~ U{30, 45, 63, 90, 128} over
    30 km, T_sfc ~ U[285, 305] K, alpha_w ~ U[1e-4, 1e-2] 1/s,
    K ~ U[0.1, 50] m^2/s, seeded.  Returns (members, initial states)."""
    rng = np.random.default_rng(seed)
    sizes = rng.choice([30, 45, 63, 90, 128], size=n)
    T_s = rng.uniform(285.0, 305.0, size=n)
    aw = rng.uniform(1e-4, 1e-2, size=n)
    K = rng.uniform(0.1, 50.0, size=n)
    members, states = [], []
    for m in range(n):
        zf = member_faces(int(sizes[m]))
        members.append(Member(zf, float(aw[m]), float(K[m]), params.f_plane))
        g = G.columns(1, zf, params.f_plane)
        states.append(I.column_state(g, params, seed=seed * 100003 + m, T_s=np.array([T_s[m]])))
    return members, states


class ColumnEnsemble:
    """Native column ensemble on the current GPU (no CPU fallback)."""

    def __init__(self, members, params: Params, numerics: Numerics, device=None, lib_path: str = LIB_PATH):
        import torch

        if not torch.cuda.is_available():
            raise RuntimeError("ColumnEnsemble needs a GPU (no CPU fallback)")
        self.torch = torch
        self.device = torch.device(device or "cuda:0")
        torch.cuda.set_device(self.device)
        self.lib = load_library(lib_path)
        self.members = list(members)
        self._faces = [np.ascontiguousarray(m.z_face, np.float64) for m in self.members]
        arr = (CaMember * len(self.members))()
        for i, (m, zf) in enumerate(zip(self.members, self._faces)):
            arr[i] = CaMember(m.nlev, zf.ctypes.data_as(C.POINTER(C.c_double)), m.alpha_w, m.D0, m.f3)
        p = CaParams(params.R_d, params.cp_d, params.cv_d, params.T_0, params.p_ref_theta, params.grav,
                     params.T_surf_ref, params.T_min_ref, params.T_min_sgs)
        c = CaConfig()
        c.hyperdiff = 0
        c.upwinding = int(numerics.upwinding)
        c.rayleigh = int(bool(numerics.rayleigh))
        c.zd_rayleigh, c.alpha_uh, c.alpha_w = params.zd_rayleigh, params.alpha_rayleigh_uh, params.alpha_rayleigh_w
        c.vert_diff = int(bool(numerics.vert_diff))
        c.vert_diff_model = numerics.vert_diff_kind()
        c.vert_diff_H = float(numerics.vert_diff_H)
        c.vert_diff_C_E = float(numerics.vert_diff_C_E)
        c.dt = numerics.dt
        h = C.c_void_p()
        st = self.lib.ca_ensemble_create(C.byref(p), C.byref(c), len(self.members), arr, C.byref(h))
        if st != CA_OK:
            msg = self.lib.ca_ensemble_last_error(None)
            raise RuntimeError(f"ca_ensemble_create failed: {_STATUS.get(st, st)}: {msg.decode() if msg else ''}")
        self.h = h
        self.S = int(self.lib.ca_ensemble_state_size(h))
        self.off = [int(self.lib.ca_ensemble_member_offset(h, i)) for i in range(len(self.members))]

    def close(self):
        if getattr(self, "h", None):
            self.lib.ca_ensemble_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, st, what):
        if st != CA_OK:
            msg = self.lib.ca_ensemble_last_error(self.h)
            raise RuntimeError(f"{what}: {_STATUS.get(st, st)}: {msg.decode() if msg else ''}")

    def _ptr(self, t):
        assert t.is_cuda and t.dtype == self.torch.float32 and t.is_contiguous() and t.numel() == self.S
        return C.c_void_p(t.data_ptr())

    def _stream(self, stream=None):
        s = stream if stream is not None else self.torch.cuda.current_stream(self.device)
        return C.c_void_p(s.cuda_stream)

    # ---------------------------------------------------------------- state
    def state(self, states):
        """Flat device state from per-member packed numpy states."""
        flat = np.concatenate([np.asarray(s, np.float32) for s in states])
        assert flat.size == self.S
        return self.torch.tensor(flat, device=self.device)

    def unpack(self, Y):
        y = Y.detach().cpu().numpy()
        return [y[o:o + 5 * m.nlev + 1] for o, m in zip(self.off, self.members)]

    # ----------------------------------------------------------------- step
    def step(self, Y, dt, nsteps=1):
        self._check(self.lib.ca_ensemble_step(self.h, self._ptr(Y), float(dt), int(nsteps), self._stream()),
                    "ca_ensemble_step")

    def step_graph(self, Y, dt, nsteps, stream):
        """The persistent launch captured in a hipGraph on ``stream`` (a
        torch.cuda.Stream other than the legacy default stream)."""
        self._check(self.lib.ca_ensemble_step_graph(self.h, self._ptr(Y), float(dt), int(nsteps),
                                                    self._stream(stream)), "ca_ensemble_step_graph")

    def set_tuning(self, lanes=None, order=None):
        if lanes is not None:
            self._check(self.lib.ca_ensemble_set_tuning(self.h, CA_ENS_LANES, int(lanes)), "set lanes")
        if order is not None:
            self._check(self.lib.ca_ensemble_set_tuning(self.h, CA_ENS_ORDER, int(order)), "set order")

    def get_tuning(self):
        v = C.c_int32()
        out = {}
        for name, knob in (("lanes", CA_ENS_LANES), ("order", CA_ENS_ORDER)):
            self._check(self.lib.ca_ensemble_get_tuning(self.h, knob, C.byref(v)), "get tuning")
            out[name] = v.value
        return out

    def set_params(self, alpha_w=None, D0=None):
        """New per-member Rayleigh coefficients and / or diffusivities."""
        def arr(x):
            if x is None:
                return None, None
            a = np.ascontiguousarray(np.asarray(x, np.float64).reshape(-1))
            assert a.size == len(self.members)
            return a, a.ctypes.data_as(C.POINTER(C.c_double))
        aw, paw = arr(alpha_w)
        d0, pd0 = arr(D0)
        self._check(self.lib.ca_ensemble_set_params(self.h, paw, pd0), "ca_ensemble_set_params")
        for i, m in enumerate(self.members):
            if aw is not None:
                m.alpha_w = float(aw[i])
            if d0 is not None:
                m.D0 = float(d0[i])

    def adjoint(self, Y0, YbN, dt, nsteps, every=8, noalloc=False):
        """Reverse mode of ``nsteps`` steps from Y0: given the cotangent YbN of
        the states after the steps, returns the cotangents of Y0 and of each
        member's D0 and alpha_w (device tensors).  ``noalloc`` calls
        ca_ensemble_adjoint_noalloc (scratch in the library's fixed per-GPU
        array instead of handle allocations)."""
        t = self.torch
        Yb0 = t.empty_like(Y0)
        M = len(self.members)
        D0b = t.empty(M, dtype=t.float32, device=self.device)
        awb = t.empty(M, dtype=t.float32, device=self.device)
        assert YbN.is_cuda and YbN.dtype == t.float32 and YbN.is_contiguous() and YbN.numel() == self.S
        fn = self.lib.ca_ensemble_adjoint_noalloc if noalloc else self.lib.ca_ensemble_adjoint
        self._check(fn(self.h, self._ptr(Y0), C.c_void_p(YbN.data_ptr()), float(dt), int(nsteps), int(every),
                       C.c_void_p(Yb0.data_ptr()), C.c_void_p(D0b.data_ptr()), C.c_void_p(awb.data_ptr()),
                       self._stream()), fn.__name__)
        return Yb0, D0b, awb

    def autotune(self, Y, dt, nsteps):
        """Times every schedule on a copy of Y and keeps the fastest; returns
        {(lanes, order): ms}."""
        t = (C.c_double * 6)()
        self._check(self.lib.ca_ensemble_autotune(self.h, self._ptr(Y), float(dt), int(nsteps), self._stream(), t),
                    "ca_ensemble_autotune")
        return {(lanes, order): t[2 * i + order] for i, lanes in enumerate((64, 128, 256)) for order in (0, 1)}
