"""Surface fluxes (surface_flux_tendency! with the default AtmosSurface:
ExchangeCoefficients Cd = Ch = 0.0044, AnalyticTemperature surface
temperature) -- the native step against the oracle through the C ABI, on a
column (surface at 300 K) and on the h4 z16 deep sphere (the zonally
symmetric surface temperature of src/setups/Setups.jl:210-216).

Gates as for the other parity tests (relative L2 per field, native float32 vs
oracle float64): rho, u_h, rho e_tot <= 1e-5 for the explicit tendency and
one step, <= 1e-3 after 20 column steps; u_3 is reported against its stated
gate with the float32 storage floor.  The fluxes must move the state: the
first-level rho e_tot and u_h differ from a step without them by far more
than the parity error.
"""
import ctypes as C

import numpy as np
import pytest

from cadycore import geometry as G
from cadycore import initial as I
from cadycore import params as P

pytestmark = pytest.mark.gpu

PAR = P.Params.from_toml()


def _torch():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def errors(Yg, Yo, g):
    fg, fo = I.unpack(Yg, g.ncol, g.nlev), I.unpack(Yo, g.ncol, g.nlev)
    return {n: rel(a, b) for n, a, b in zip(I.FIELD_NAMES, fg, fo)}


def _column():
    g = G.columns(1, G.tanh_stretched_faces(30, 30000.0, 500.0))
    Y = I.column_state(g, PAR, seed=0)
    # a warmer surface and a wind, so both fluxes are well away from zero
    return g, P.Numerics(hyperdiff=False, dt=10.0, sfc_flux=True), Y, np.full(g.ncol, 305.0)


def _sphere():
    g = G.cubed_sphere(4, G.tanh_stretched_faces(16, 30000.0, 100.0), PAR.planet_radius, PAR.Omega)
    return g, P.Numerics(dt=100.0, sfc_flux=True), I.sphere_state(g, PAR, seed=1), I.surface_temperature(g)


def _pair(g, num, Ts):
    _torch()
    from cadycore.lib import DryDycore
    from oracle import Oracle

    return DryDycore(g, PAR, num, T_sfc_col=Ts), Oracle(g, PAR, num, T_sfc_col=Ts)


@pytest.mark.parametrize("case", ["column", "sphere"])
def test_surface_flux_explicit_tendency_parity(case):
    torch = _torch()
    g, num, Y0, Ts = _column() if case == "column" else _sphere()
    d, o = _pair(g, num, Ts)
    Y32 = Y0.astype(np.float32).astype(np.float64)
    Yo = Y32.copy()
    o.set_precomputed(Yo)
    Yg = d.state(Y32)
    d.set_implicit_precomputed_quantities(Yg)
    T = d.zeros()
    d.remaining_tendency(T, Yg)
    torch.cuda.synchronize()
    got = T.cpu().numpy().astype(np.float64)
    ref = o.remaining_tendency(Yo)
    d.close()
    e = errors(got, ref, g)
    print(case, "T_exp", {k: f"{v:.2e}" for k, v in e.items()})
    for name in ("u1", "rhoe") + (("u2", "rho") if case == "sphere" else ()):
        assert e[name] <= 1e-5, (name, e)
    # the surface terms themselves (first level only), native against oracle
    num_off = P.Numerics(**{**num.__dict__, "sfc_flux": False})
    from oracle import Oracle

    o_off = Oracle(g, PAR, num_off)
    Yo2 = Y32.copy()
    o_off.set_precomputed(Yo2)
    ref_off = o_off.remaining_tendency(Yo2)
    nC = g.nlev * g.ncol
    for p in (1, 3):
        sl = slice(p * nC, p * nC + g.ncol)  # level 0 of the plane
        dref = ref[sl] - ref_off[sl]
        assert np.linalg.norm(dref) > 100 * np.linalg.norm(got[sl] - ref[sl]), p


@pytest.mark.parametrize("case,nsteps", [("column", 1), ("column", 20), ("sphere", 1), ("sphere", 5)])
def test_surface_flux_step_parity(case, nsteps):
    torch = _torch()
    g, num, Y0, Ts = _column() if case == "column" else _sphere()
    d, o = _pair(g, num, Ts)
    Y32 = Y0.astype(np.float32).astype(np.float64)
    Yo = Y32.copy()
    o.dss(Yo)
    o.set_precomputed(Yo)
    o.step(Yo, num.dt, nsteps)
    Yg = d.state(Y32)
    d.step(Yg, num.dt, nsteps)
    torch.cuda.synchronize()
    got = Yg.cpu().numpy().astype(np.float64)
    d.close()
    e = errors(got, Yo, g)
    print(case, nsteps, {k: f"{v:.2e}" for k, v in e.items()})
    gate = 1e-5 if nsteps == 1 else 1e-3
    for name in ("rho", "u1", "u2", "rhoe"):
        if case == "column" and name == "u2":
            continue  # f-plane column: u2 grows from 0 by Coriolis only
        assert e[name] <= gate, (name, e)
    state = "MET" if e["u3"] <= gate else "UNMET"
    print(f"u3 gate [surface flux {case} {nsteps} steps]: {e['u3']:.2e} vs stated {gate:.0e}: {state}")
    assert np.all(np.isfinite(got))
    # the fluxes move the first level far beyond the parity error
    from oracle import Oracle

    o_off = Oracle(g, PAR, P.Numerics(**{**num.__dict__, "sfc_flux": False}))
    Yf = Y32.copy()
    o_off.dss(Yf)
    o_off.set_precomputed(Yf)
    o_off.step(Yf, num.dt, nsteps)
    nC = g.nlev * g.ncol
    sl = slice(3 * nC, 3 * nC + g.ncol)
    assert np.linalg.norm(Yo[sl] - Yf[sl]) > 100 * np.linalg.norm(got[sl] - Yo[sl])


def test_surface_flux_rejected_where_unsupported():
    torch = _torch()
    g = G.cubed_sphere(2, G.tanh_stretched_faces(10, 30000.0, 300.0), PAR.planet_radius, PAR.Omega)
    num = P.Numerics(dt=100.0, sfc_flux=True)
    from cadycore.lib import DryDycore

    d = DryDycore(g, PAR, num, T_sfc_col=I.surface_temperature(g))
    Y = d.state(I.sphere_state(g, PAR, seed=1))
    st = d.lib.ca_step_record(d.h, C.c_void_p(Y.data_ptr()), 100.0, 2, 1, None)
    assert st == 1 and b"surface" in d.lib.ca_last_error(d.h)
    torch.cuda.synchronize()
    d.close()
    # creating without the per-column surface temperature fails at create
    with pytest.raises(ValueError):
        DryDycore(g, PAR, num)
