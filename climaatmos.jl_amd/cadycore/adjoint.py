"""Differentiable solves as ``torch.autograd.Function``s whose backward is the
native reverse mode:

* ``EnsembleSolve``: column ensembles, ``Y_N = step^nsteps(Y_0; D0, alpha_w)``
  (``ca_ensemble_adjoint``: checkpointing every ``every`` steps, transposed
  Schur solve, reverse ARS343 stage sweep);
* ``StepSolve``: the sphere, ``Y_N = step^nsteps(Y_0; D0)`` (forward
  ``ca_step_record`` keeps checkpoints, backward ``ca_step_pullback``).

They are the Python counterparts of ChainRulesCore ``rrule``s around the same
C entry points; the reference itself has no adjoint (it calibrates without
gradients, ``calibration/``).

``column_temperature`` and ``column_loss`` restate the calibration loss of the
workload: the mean over columns of (integral of T dz - target)^2, with T from
the dry equation of state of the final state.
"""
from __future__ import annotations

import numpy as np

from .params import Params


def _torch():
    import torch

    return torch


class EnsembleSolve:
    """Namespace for the autograd function (built lazily: torch is optional
    at import time)."""

    _fn = None

    @classmethod
    def apply(cls, Y0, D0, alpha_w, ens, dt, nsteps, every=8):
        if cls._fn is None:
            torch = _torch()

            class _Solve(torch.autograd.Function):
                @staticmethod
                def forward(ctx, Y0, D0, alpha_w, ens, dt, nsteps, every):
                    ens.set_params(alpha_w=alpha_w.detach().cpu().double().numpy(),
                                   D0=D0.detach().cpu().double().numpy())
                    Y = Y0.detach().clone().contiguous()
                    ens.step(Y, dt, nsteps)
                    ctx.save_for_backward(Y0.detach().clone())
                    ctx.ens, ctx.dt, ctx.nsteps, ctx.every = ens, dt, nsteps, every
                    ctx.pdtype = (D0.dtype, D0.device, alpha_w.dtype, alpha_w.device)
                    return Y

                @staticmethod
                def backward(ctx, Ybar):
                    (Y0,) = ctx.saved_tensors
                    Yb0, D0b, awb = ctx.ens.adjoint(Y0, Ybar.contiguous().float(), ctx.dt, ctx.nsteps, ctx.every)
                    dd, ddev, ad, adev = ctx.pdtype
                    return Yb0, D0b.to(dtype=dd, device=ddev), awb.to(dtype=ad, device=adev), None, None, None, None

            cls._fn = _Solve
        return cls._fn.apply(Y0, D0, alpha_w, ens, dt, nsteps, every)


class StepSolve:
    """``Y_N = StepSolve.apply(Y0, D0, dycore, dt, nsteps, every)`` on a
    sphere ``cadycore.lib.DryDycore`` created with vertical diffusion; D0 is
    the per-column diffusivity (any float dtype/device, copied to the
    dycore).  Differentiable in Y0 and D0."""

    _fn = None

    @classmethod
    def apply(cls, Y0, D0, dycore, dt, nsteps, every=8):
        if cls._fn is None:
            torch = _torch()

            class _Solve(torch.autograd.Function):
                @staticmethod
                def forward(ctx, Y0, D0, dycore, dt, nsteps, every):
                    dycore.set_column_params(D0_col=D0.detach().cpu().double().numpy())
                    Y = Y0.detach().clone().contiguous()
                    dycore.step_record(Y, dt, nsteps, every)
                    ctx.dycore = dycore
                    ctx.pdtype = (D0.dtype, D0.device)
                    return Y

                @staticmethod
                def backward(ctx, Ybar):
                    t = _torch()
                    d = ctx.dycore
                    D0b = t.zeros(d.grid.ncol, dtype=t.float32, device=Ybar.device)
                    Yb0 = d.step_pullback(Ybar.contiguous().float(), D0b)
                    dd, ddev = ctx.pdtype
                    return Yb0, D0b.to(dtype=dd, device=ddev), None, None, None, None

            cls._fn = _Solve
        return cls._fn.apply(Y0, D0, dycore, dt, nsteps, every)


def column_temperature(Y, members, params: Params):
    """T of every centre of every member from the flat state (torch, any
    dtype): T = T_0 + (rho e_tot / rho - K - Phi + R_d T_0) / cv_d, floored at
    T_min_sgs; K = |u_h|^2 / 2 + the centre mean of u_3^2 g^33 / 2."""
    torch = _torch()
    out, off = [], 0
    for m in members:
        nl = m.nlev
        y = Y[off:off + 5 * nl + 1]
        off += 5 * nl + 1
        rho, u1, u2, rhoe, u3 = y[:nl], y[nl:2 * nl], y[2 * nl:3 * nl], y[3 * nl:4 * nl], y[4 * nl:]
        zf = np.asarray(m.z_face, np.float64)
        zc = 0.5 * (zf[1:] + zf[:-1])
        jvf = np.empty(nl + 1)
        jvf[1:-1] = zc[1:] - zc[:-1]
        jvf[0], jvf[-1] = zf[1] - zf[0], zf[-1] - zf[-2]
        g33 = torch.tensor(1.0 / jvf ** 2, dtype=Y.dtype, device=Y.device)
        u3i = torch.cat([u3[:1] * 0, u3[1:-1], u3[-1:] * 0])
        Kv = 0.25 * (u3i[:-1] ** 2 * g33[:-1] + u3i[1:] ** 2 * g33[1:])
        K = 0.5 * (u1 ** 2 + u2 ** 2) + Kv
        phi = torch.tensor(params.grav * zc, dtype=Y.dtype, device=Y.device)
        T = params.T_0 + (rhoe / rho - K - phi + params.R_d * params.T_0) / params.cv_d
        out.append(torch.clamp(T, min=params.T_min_sgs))
    return out


def column_loss(Y, members, params: Params, target):
    """mean over members of (sum_k T_k dz_k - target_m)^2"""
    torch = _torch()
    Ts = column_temperature(Y, members, params)
    terms = []
    for T, m, t in zip(Ts, members, target):
        dz = torch.tensor(np.diff(np.asarray(m.z_face, np.float64)), dtype=Y.dtype, device=Y.device)
        terms.append((torch.sum(T * dz) - t) ** 2)
    return torch.stack(terms).mean()
