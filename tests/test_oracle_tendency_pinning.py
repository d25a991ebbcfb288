"""Pins the oracle's explicit tendency (remaining_tendency!) against what the
mathematics fixes, term by term, so that a wrong sign or a dropped term in the
restatement fails here.

* Horizontal dynamics and vertical advection
  (src/prognostic_equations/advection.jl:36-91, :205-290): manufactured smooth
  states on the deep cubed sphere; the oracle's DSSed tendency converges to the
  analytic continuous tendency -- third order in the element size for the
  horizontal terms, second order in the layer thickness for the vertical ones.
* Hyperdiffusion (hyperdiffusion.jl:116-147 prep, :247-307 apply): degree-3
  spherical harmonics are eigenfunctions; the mass-weighted decay rate equals
  nu_v (l(l+1)/r^2)^2 for a rotational wind, delta times that for a divergent
  wind, nu_s for the scalar (with rho inside the outer divergence) and nu_v for
  u_3.
* Vertical diffusion (vertical_diffusion_boundary_layer.jl:63-103): exact
  discrete eigenmodes of the zero-flux second difference, the exact product
  rule for linear rho and s_d, column conservation, and second-order
  convergence on the deep sphere.
* Rayleigh sponge on u_h (rayleigh_sponge.jl:39-44): -alpha zeta(z) u_h.
* Surface fluxes (surface_flux.jl:101-145 with ExchangeCoefficients): only the
  first centre changes; an isothermal surface layer gives the exact
  hydrostatic surface density; a surface at the interior's dry static energy
  gives no heat flux; the drag is -(rho_s / rho) Cd U u / dz with the wind
  speed floored at the gustiness; on the deep sphere the column energy gain
  equals the bulk heat flux through the surface area and the covariant drag
  is parallel to the covariant wind.
* Implicit vertical diffusion (diff_mode Implicit: implicit_tendency.jl:73-82,
  the blocks of manual_sparse_jacobian.jl:1031-1261 and the approximate
  arrowhead solve of :538-578): the tendency moves unchanged from the explicit
  to the implicit tendency; the (rho e_tot, rho e_tot) and (u_h, u_h) blocks
  equal central differences of the implicit tendency; the stationary
  iterations of the linear solve converge to the dense solve of the assembled
  column matrix, and without diffusion one iteration is that solve exactly.

Each check is a function of an Oracle factory so that
tests/test_oracle_mutants.py can run the same checks on copies of the oracle
with a planted mistake.
"""
import math

import numpy as np
import pytest

import manufactured as M
from cadycore import geometry as G
from cadycore import params as P
from oracle import Oracle

PAR = P.Params.from_toml()


def default_factory(grid, params, numerics, **kw):
    return Oracle(grid, params, numerics, **kw)


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _tendency(mk, grid, numerics, Y, dss=True, **kw):
    o = mk(grid, PAR, numerics, **kw)
    Y = Y.copy()
    o.set_precomputed(Y)
    Yt = o.remaining_tendency(Y)
    if dss:
        o.dss(Yt)
    return Yt


def _planes(grid, Yt):
    nl, nc = grid.nlev, grid.ncol
    nC = nl * nc
    return [Yt[i * nC:(i + 1) * nC].reshape(nl, nc) for i in range(4)] + [Yt[4 * nC:].reshape(nl + 1, nc)]


def _order(errs):
    return [math.log2(errs[i] / errs[i + 1]) for i in range(len(errs) - 1)]


# ------------------------------------------------- manufactured solutions

def mms_horizontal(mk, h_elems=(3, 6, 12), nlev=4):
    """Relative L2 errors [(rho, u1, u2, rhoe)] of the DSSed explicit tendency
    against the continuous one, w = 0, hyperdiffusion off."""
    sm = M.Smooth(PAR)
    out = []
    for h in h_elems:
        g = G.cubed_sphere(h, G.uniform_faces(nlev, 30000.0), PAR.planet_radius, PAR.Omega)
        got = _planes(g, _tendency(mk, g, P.Numerics(dt=100.0, hyperdiff=False), sm.state(g)))
        ana = sm.explicit_tendency(g)
        out.append([_rel(got[i], ana[i]) for i in range(4)])
    return out


def mms_vertical(mk, nlevs=(8, 16, 32), h_elem=4, w_amp=0.0):
    """Relative L2 errors of the explicit u_3 tendency at interior faces under
    vertical refinement (w = w_amp profile)."""
    sm = M.Smooth(PAR, w_amp=w_amp)
    out = []
    for nl in nlevs:
        g = G.cubed_sphere(h_elem, G.uniform_faces(nl, 30000.0), PAR.planet_radius, PAR.Omega)
        got = _planes(g, _tendency(mk, g, P.Numerics(dt=100.0, hyperdiff=False), sm.state(g)))
        ana = sm.explicit_tendency(g)
        out.append(_rel(got[4][1:-1], ana[4][1:-1]))
    return out


def mms_with_w(mk, pairs=((3, 8), (6, 16), (12, 32)), w_amp=2.0):
    """Relative L2 errors [(u1, u2, u3)] with a vertical velocity, refining the
    elements and the layers together."""
    sm = M.Smooth(PAR, w_amp=w_amp)
    out = []
    for h, nl in pairs:
        g = G.cubed_sphere(h, G.uniform_faces(nl, 30000.0), PAR.planet_radius, PAR.Omega)
        got = _planes(g, _tendency(mk, g, P.Numerics(dt=100.0, hyperdiff=False), sm.state(g)))
        ana = sm.explicit_tendency(g)
        out.append([_rel(got[1], ana[1]), _rel(got[2], ana[2]), _rel(got[4][1:-1], ana[4][1:-1])])
    return out


def check_mms_horizontal(mk):
    e = np.array(mms_horizontal(mk))
    for f in range(4):
        order = _order(e[:, f])
        assert min(order) > 2.7, f"field {f}: orders {order} errors {e[:, f]}"
    assert e[-1, 0] < 3e-4 and e[-1, 3] < 3e-4, e[-1]
    assert e[-1, 1] < 4e-5 and e[-1, 2] < 4e-5, e[-1]


def check_mms_vertical(mk):
    e = mms_vertical(mk)
    assert min(_order(e)) > 1.9 and e[-1] < 1.5e-5, e


def check_mms_with_w(mk):
    e = np.array(mms_with_w(mk))
    for f in range(3):
        order = _order(e[:, f])
        assert min(order) > 1.8, f"field {f}: orders {order} errors {e[:, f]}"
    assert e[-1, 0] < 3e-3 and e[-1, 1] < 1e-3 and e[-1, 2] < 1.5e-4, e[-1]


def test_mms_horizontal_dynamics_third_order():
    check_mms_horizontal(default_factory)


def test_mms_vertical_advection_second_order():
    check_mms_vertical(default_factory)


def test_mms_with_vertical_velocity_converges():
    check_mms_with_w(default_factory)


# -------------------------------------------------------- hyperdiffusion

def hyperdiffusion_rates(mk, h_elem, nlev=6):
    """Mass-weighted decay rate of each harmonic state's hyperdiffusion
    tendency relative to the closed form (1 = exact)."""
    _, w, _ = G.gll()
    g = G.cubed_sphere(h_elem, G.uniform_faces(nlev, 30000.0), PAR.planet_radius, PAR.Omega)
    WJ = np.tile(np.outer(w, w).reshape(-1), g.nelem) * g.J_h
    on, off = P.Numerics(dt=100.0, hyperdiff=True), P.Numerics(dt=100.0, hyperdiff=False)
    nu_v, nu_s = on.nu4(g.h)
    rates = {}
    for kind in ("rot", "div", "scalar", "w"):
        Y, info = M.harmonic_state(g, PAR, kind)
        d = _planes(g, _tendency(mk, g, on, Y) - _tendency(mk, g, off, Y))
        exp = M.hyperdiffusion_closed_form(g, PAR, kind, info, nu_v, nu_s, on.div_damping)
        if kind in ("rot", "div"):
            u = (info["u1"], info["u2"])
            rates[kind] = sum((WJ * d[1 + i] * u[i]).sum() for i in range(2)) / \
                sum((WJ * exp[i] * u[i]).sum() for i in range(2))
        elif kind == "scalar":
            rates[kind] = (WJ * d[3] * exp).sum() / (WJ * exp * exp).sum()
        else:
            rates[kind] = (WJ * d[4][2:-2] * exp[2:-2]).sum() / (WJ * exp[2:-2] ** 2).sum()
    return rates


def check_hyperdiffusion(mk):
    r8, r16 = hyperdiffusion_rates(mk, 8), hyperdiffusion_rates(mk, 16)
    for kind, tol in (("rot", 1.5e-3), ("div", 3e-4), ("scalar", 1e-4), ("w", 1e-3)):
        assert abs(r16[kind] - 1) < tol, (kind, r8[kind], r16[kind])
        assert abs(r16[kind] - 1) <= abs(r8[kind] - 1) + 1e-5, (kind, r8[kind], r16[kind])


def test_hyperdiffusion_eigenmode_decay_rates():
    check_hyperdiffusion(default_factory)


# ----------------------------------------------------- vertical diffusion

def _column_state(zc, rho, T, u1):
    K = 0.5 * u1 ** 2
    rhoe = rho * (PAR.cv_d * (T - PAR.T_0) - PAR.R_d * PAR.T_0 + K + PAR.grav * zc)
    return np.concatenate([rho, u1, 0 * u1, rhoe, np.zeros(len(zc) + 1)])


def check_vertical_diffusion_column(mk):
    N, H, D0 = 12, 3000.0, 7.0
    g = G.columns(1, G.uniform_faces(N, H))
    zc, dz = g.zc, H / N
    on = P.Numerics(dt=10.0, hyperdiff=False, vert_diff=True)
    off = P.Numerics(dt=10.0, hyperdiff=False)
    D = np.array([D0])

    def vd(Y):
        return _tendency(mk, g, on, Y, dss=False, D0_col=D) - _tendency(mk, g, off, Y, dss=False)

    m = 3
    cosm = np.cos(np.pi * m * (np.arange(N) + 0.5) / N)
    lam = 4.0 / dz ** 2 * math.sin(math.pi * m / (2 * N)) ** 2
    # s_d = cp_d (T - T_0) + g z is a cosine mode of the zero-flux second difference
    T = 300.0 - PAR.grav / PAR.cp_d * zc + 2.0 * cosm
    d = vd(_column_state(zc, np.full(N, 1.1), T, np.zeros(N)))
    exp = -D0 * lam * 1.1 * PAR.cp_d * 2.0 * cosm
    assert np.abs(d[3 * N:4 * N] - exp).max() < 1e-10 * np.abs(exp).max()
    # u_h cosine mode
    u = 5.0 * cosm
    d = vd(_column_state(zc, np.full(N, 1.1), np.full(N, 280.0), u))
    assert np.abs(d[N:2 * N] + D0 * lam * u).max() < 1e-10 * np.abs(D0 * lam * u).max()
    # linear rho and linear s_d: interior rho e_t = D0 (d rho/dz)(d s_d/dz) exactly
    rho, T = 1.2 - 1e-4 * zc, 300.0 - 0.003 * zc
    d = vd(_column_state(zc, rho, T, np.zeros(N)))
    exp = D0 * (-1e-4) * (PAR.cp_d * (-0.003) + PAR.grav)
    np.testing.assert_allclose(d[3 * N + 1:4 * N - 1], exp, rtol=1e-9)
    # zero-flux boundaries: column energy and momentum conserved
    rng = np.random.default_rng(0)
    Y = _column_state(zc, 1 + 0.1 * rng.random(N), 280 + 5 * rng.random(N), 3 * rng.standard_normal(N))
    d = vd(Y)
    scale = np.abs(d[3 * N:4 * N]).max()
    assert abs(d[3 * N:4 * N].sum()) < 1e-12 * scale * N
    assert abs((d[N:2 * N] * Y[:N]).sum()) < 1e-12 * np.abs(d[N:2 * N] * Y[:N]).max() * N


def vertical_diffusion_sphere_errors(mk, nlevs=(8, 16, 32), D0=20.0):
    sm = M.Smooth(PAR)
    out = []
    for nl in nlevs:
        g = G.cubed_sphere(2, G.uniform_faces(nl, 30000.0), PAR.planet_radius, PAR.Omega)
        on = P.Numerics(dt=100.0, hyperdiff=False, vert_diff=True)
        off = P.Numerics(dt=100.0, hyperdiff=False)
        Y = sm.state(g)
        d = _planes(g, _tendency(mk, g, on, Y, dss=False, D0_col=np.full(g.ncol, D0))
                    - _tendency(mk, g, off, Y, dss=False))
        ex = M.vdiff_tendency(sm, g, D0)
        out.append([_rel(d[3][1:-1], ex[0][1:-1]), _rel(d[1][1:-1], ex[1][1:-1]),
                    _rel(d[2][1:-1], ex[2][1:-1])])
    return out


def check_vertical_diffusion_sphere(mk):
    e = np.array(vertical_diffusion_sphere_errors(mk))
    for f in range(3):
        assert min(_order(e[:, f])) > 1.9 and e[-1, f] < 3e-3, e


def check_vertical_diffusion_closures(mk):
    """The two closures of ᶜcompute_eddy_diffusivity_coefficient
    (eddy_diffusivity_coefficient.jl:16-41), with the face diffusivity the
    harmonic mean interp(rho) / interp(1 / K) (vertical_diffusion_boundary_layer.jl:86-90).
    On uniform layers with constant rho and linear s_d the interior tendency is
    rho s_d' (K_{k+1/2} - K_{k-1/2}) / dz exactly, which fixes the face K:
    * DecayWithHeightDiffusion, K = D0 exp(-z / H): the harmonic face mean is
      D0 exp(-z_f / H) / cosh(dz / 2H) (an arithmetic mean would multiply by
      cosh instead);
    * VerticalDiffusion, K = C_E |u_h(lowest centre)| dz_1 / 2, times
      exp(-((85000 - p) / 10000)^2) where p < 85000 Pa: uniform below 850 hPa,
      where the step equals DecayWithHeightDiffusion with D0 = C_E |u| dz / 2."""
    N, Htop = 16, 4000.0
    g = G.columns(1, G.uniform_faces(N, Htop))
    zc, dz, zf = g.zc, Htop / N, g.z_face
    off = P.Numerics(dt=10.0, hyperdiff=False)
    rho0, Tb, Tt = 1.1, 300.0, 240.0
    T = Tb + (Tt - Tb) * zc / Htop
    sp = PAR.cp_d * (Tt - Tb) / Htop + PAR.grav  # d s_d / dz
    Y = _column_state(zc, np.full(N, rho0), T, np.full(N, 5.0))

    def vd(num, **kw):
        return (_tendency(mk, g, num, Y, dss=False, **kw) - _tendency(mk, g, off, Y, dss=False))[3 * N:4 * N]

    # DecayWithHeightDiffusion with H = 800 m
    D0, H = 6.0, 800.0
    dec = P.Numerics(dt=10.0, hyperdiff=False, vert_diff=True, vert_diff_H=H)
    Kf = D0 * np.exp(-zf / H) / math.cosh(dz / (2 * H))
    exp = rho0 * sp * (Kf[2:N] - Kf[1:N - 1]) / dz
    np.testing.assert_allclose(vd(dec, D0_col=np.array([D0]))[1:-1], exp, rtol=1e-9)
    # VerticalDiffusion: p = rho R_d T crosses 850 hPa inside the column
    CE = 0.044
    bulk = P.Numerics(dt=10.0, hyperdiff=False, vert_diff=True, vert_diff_model="VerticalDiffusion", vert_diff_C_E=CE)
    p = rho0 * PAR.R_d * T
    assert p[0] > 85000 > p[-1]
    Kc = CE * 5.0 * dz / 2 * np.where(p > 85000, 1.0, np.exp(-((85000 - p) / 10000) ** 2))
    Kf = 2 / (1 / Kc[:-1] + 1 / Kc[1:])  # interior faces 1 .. N - 1
    exp = rho0 * sp * (Kf[1:] - Kf[:-1]) / dz
    np.testing.assert_allclose(vd(bulk)[1:-1], exp, rtol=1e-9)
    # below 850 hPa the closure is a uniform K = C_E |u| dz / 2
    g2 = G.columns(1, G.uniform_faces(N, 800.0))
    Y2 = _column_state(g2.zc, np.full(N, rho0), 300.0 - 0.006 * g2.zc, np.full(N, 5.0))
    a = _tendency(mk, g2, bulk, Y2, dss=False)
    b = _tendency(mk, g2, P.Numerics(dt=10.0, hyperdiff=False, vert_diff=True), Y2, dss=False,
                  D0_col=np.array([CE * 5.0 * (800.0 / N) / 2]))
    assert np.abs(a - b).max() <= 1e-12 * np.abs(b).max()


def test_vertical_diffusion_column_closed_forms():
    check_vertical_diffusion_column(default_factory)


def test_vertical_diffusion_closures():
    check_vertical_diffusion_closures(default_factory)


def test_vertical_diffusion_deep_sphere_second_order():
    check_vertical_diffusion_sphere(default_factory)


# --------------------------------------------------------------- Rayleigh

def check_rayleigh_uh(mk):
    zf = G.uniform_faces(10, 30000.0)
    g = G.cubed_sphere(2, zf, PAR.planet_radius, PAR.Omega)
    Y = M.Smooth(PAR).state(g)
    alpha = PAR.alpha_rayleigh_uh if PAR.alpha_rayleigh_uh != 0 else 1e-3

    class _P:  # the same constants with a nonzero u_h coefficient
        pass

    par = _P()
    par.__dict__.update(PAR.__dict__)
    par.alpha_rayleigh_uh = alpha
    on = P.Numerics(dt=100.0, hyperdiff=False, rayleigh=True)
    off = P.Numerics(dt=100.0, hyperdiff=False)
    o_on, o_off = mk(g, par, on), mk(g, par, off)
    a, b = Y.copy(), Y.copy()
    o_on.set_precomputed(a)
    o_off.set_precomputed(b)
    d = _planes(g, o_on.remaining_tendency(a) - o_off.remaining_tendency(b))
    u = _planes(g, Y)
    zmax, zd = zf[-1], PAR.zd_rayleigh
    zeta = np.array([math.sin(math.pi / 2 * (z - zd) / (zmax - zd)) ** 2 if z > zd else 0.0 for z in g.zc])
    for i in (1, 2):
        exp = -alpha * zeta[:, None] * u[i]
        assert np.abs(d[i] - exp).max() <= 1e-12 * np.abs(exp).max() + 1e-300, i
    assert np.abs(d[0]).max() == 0 and np.abs(d[3]).max() == 0


def test_rayleigh_uh_closed_form():
    check_rayleigh_uh(default_factory)


# ---------------------------------------------------------- surface fluxes

def _sfc(mk, grid, Y, T_sfc, Cd=0.0044, Ch=0.0044, gust=1.0):
    """The surface-flux part of the oracle's tendency (fluxes on minus off)."""
    on = P.Numerics(dt=10.0, hyperdiff=False, sfc_flux=True, Cd=Cd, Ch=Ch, gustiness=gust)
    off = P.Numerics(dt=10.0, hyperdiff=False)
    return (_tendency(mk, grid, on, Y, dss=False, T_sfc_col=np.asarray(T_sfc, float))
            - _tendency(mk, grid, off, Y, dss=False))


def check_surface_flux_column(mk):
    N, H = 10, 3000.0
    g = G.columns(1, G.uniform_faces(N, H))
    zc, dz1, z1 = g.zc, H / N, g.zc[0]
    R, cp, grav = PAR.R_d, PAR.cp_d, PAR.grav
    rho = 1.2 * np.exp(-zc / 8000.0)
    T = 290.0 - 0.0065 * zc
    for speed in (5.0, 0.3):  # above and below the 1 m/s gustiness floor
        U = max(speed, 1.0)
        # isothermal surface layer: rho_s is the exact hydrostatic extrapolation
        Y = _column_state(zc, rho, T, np.full(N, speed))
        d = _sfc(mk, g, Y, [T[0]])
        rho_s = rho[0] * math.exp(grav * z1 / (R * T[0]))
        exp_e = -rho_s * 0.0044 * U * grav * z1 / dz1
        assert abs(d[3 * N] - exp_e) < 1e-12 * abs(exp_e), (speed, d[3 * N], exp_e)
        exp_u = -rho_s / rho[0] * 0.0044 * U * speed / dz1
        assert abs(d[N] - exp_u) < 1e-12 * abs(exp_u), (speed, d[N], exp_u)
        # only the first centre of rho e_tot and u_h changes
        assert np.abs(d[:N]).max() == 0 and np.abs(d[4 * N:]).max() == 0
        assert np.abs(d[N + 1:2 * N]).max() == 0 and np.abs(d[3 * N + 1:4 * N]).max() == 0
    # surface at the interior's dry static energy: no heat flux; drag unchanged
    Y = _column_state(zc, rho, T, np.full(N, 5.0))
    Ts = T[0] + grav * z1 / cp
    d = _sfc(mk, g, Y, [Ts])
    assert abs(d[3 * N]) < 1e-12 * rho[0] * 0.0044 * 5.0 * cp * 10.0 / dz1
    # a warmer surface heats the air, linearly in Ch; Cd does not enter it
    d1 = _sfc(mk, g, Y, [Ts + 10.0])
    d2 = _sfc(mk, g, Y, [Ts + 10.0], Ch=0.0088, Cd=0.001)
    assert d1[3 * N] > 0 and abs(d2[3 * N] / d1[3 * N] - 2.0) < 1e-12
    assert abs(d2[N] / d1[N] - 0.001 / 0.0044) < 1e-12


def check_surface_flux_sphere(mk):
    zf = G.uniform_faces(6, 3000.0)
    g = G.cubed_sphere(2, zf, PAR.planet_radius, PAR.Omega)
    Y = M.Smooth(PAR).state(g)
    rng = np.random.default_rng(5)
    T_sfc = 285.0 + 10.0 * rng.random(g.ncol)
    d = _planes(g, _sfc(mk, g, Y, T_sfc))
    u = _planes(g, Y)
    assert all(np.abs(d[i][1:]).max() == 0 for i in (1, 2, 3)) and np.abs(d[0]).max() == 0
    # the covariant drag is parallel to the covariant wind (the metric round trip)
    r1, r2 = d[1][0] / u[1][0], d[2][0] / u[2][0]
    big = (np.abs(u[1][0]) > 1.0) & (np.abs(u[2][0]) > 1.0)
    assert big.sum() > 10 and np.abs(r1[big] / r2[big] - 1).max() < 1e-10
    assert np.all(r1[big] < 0)
    # energy: the column gain equals the bulk heat flux through the surface
    # area; the wind speed is |(u^1 |a_1|, u^2 |a_2|)| at the first centre
    a = PAR.planet_radius
    s1, s0 = (a + g.zc[0]) / a, (a + zf[0]) / a
    gu, gd = g.gup, g.gdn
    ct1 = (gu[0] * u[1][0] + gu[1] * u[2][0]) / s1 ** 2
    ct2 = (gu[1] * u[1][0] + gu[2] * u[2][0]) / s1 ** 2
    U = np.maximum(np.hypot(ct1 * np.sqrt(gd[0]) * s1, ct2 * np.sqrt(gd[2]) * s1), 1.0)
    o = mk(g, PAR, P.Numerics(dt=10.0, hyperdiff=False))
    y = Y.copy()
    o.set_precomputed(y)
    T = o.precomputed()[1].reshape(-1)
    T1, rho1, z1 = T[:g.ncol], u[0][0], g.zc[0]
    rho_s = rho1 * T1 / T_sfc * np.exp(PAR.grav * z1 / (PAR.R_d * (T1 + T_sfc) / 2))
    shf = -rho_s * 0.0044 * U * (PAR.cp_d * (T1 - T_sfc) + PAR.grav * z1)
    Jc1 = g.J_h * s1 ** 2 * (zf[1] - zf[0])
    np.testing.assert_allclose(d[3][0] * Jc1, g.J_h * s0 ** 2 * shf, rtol=1e-10)
    # drag magnitude from the same physics: rho_s Cd U through the surface area
    # over the cell volume.  The flux's covariant components are formed with
    # the surface metric and added to the centre's unchanged, so the physical
    # drag at the centre is s_0 / s_1 (1 - 4e-5 here) of the surface one.
    np.testing.assert_allclose(-r1, rho_s / rho1 * 0.0044 * U * g.J_h * s0 ** 2 / Jc1 * (s0 / s1), rtol=1e-10)


def test_surface_flux_column_closed_forms():
    check_surface_flux_column(default_factory)


def test_surface_flux_deep_sphere_budget():
    check_surface_flux_sphere(default_factory)


# ------------------------------------------------ implicit vertical diffusion

def _num_diff(implicit, iters=1, **kw):
    return P.Numerics(dt=100.0, hyperdiff=False, vert_diff=True, implicit_diffusion=implicit, solve_iters=iters, **kw)


def check_implicit_diffusion_moves_tendency(mk):
    g = G.cubed_sphere(2, G.uniform_faces(8, 30000.0), PAR.planet_radius, PAR.Omega)
    Y = M.Smooth(PAR).state(g)
    D0 = np.random.default_rng(2).uniform(0.5, 50.0, g.ncol)

    def both(implicit):
        o = mk(g, PAR, _num_diff(implicit), D0_col=D0)
        y = Y.copy()
        o.set_precomputed(y)
        return o.remaining_tendency(y), o.implicit_tendency(y)

    te, ti = both(False)
    ie, ii = both(True)
    moved = te - ie
    assert np.linalg.norm(moved) > 0
    np.testing.assert_allclose(ii - ti, moved, rtol=0, atol=1e-12 * np.abs(moved).max())


def _dense_column(b, nl):
    """The assembled Jacobian of one column from its blocks, ordered
    [rho | u_1 | u_2 | rho e_tot | u_3]."""
    n = 5 * nl + 1
    A = np.zeros((n, n))
    r, h1, h2, e, w = (np.arange(nl), nl + np.arange(nl), 2 * nl + np.arange(nl), 3 * nl + np.arange(nl),
                       4 * nl + np.arange(nl + 1))
    A[r, r] = -1.0
    for k in range(nl):
        for blk, rows in (("Aee", e), ("Ahh", h1), ("Ahh", h2)):
            lo, d, up = b[blk][k]
            A[rows[k], rows[k]] = d
            if k >= 1:
                A[rows[k], rows[k - 1]] = lo
            if k + 1 < nl:
                A[rows[k], rows[k + 1]] = up
        A[r[k], w[k]], A[r[k], w[k + 1]] = b["Aru"][k]
        A[e[k], w[k]], A[e[k], w[k + 1]] = b["Aeu"][k]
    for f in range(nl + 1):
        A[w[f], w[f]] = b["di"][f]
        if f >= 1:
            A[w[f], w[f - 1]] = b["lo"][f]
        if f + 1 <= nl:
            A[w[f], w[f + 1]] = b["up"][f]
        if 1 <= f <= nl - 1:
            A[w[f], r[f - 1]], A[w[f], r[f]] = b["Aur"][f]
            A[w[f], e[f - 1]], A[w[f], e[f]] = b["Aue"][f]
            A[w[f], h1[f - 1]], A[w[f], h2[f - 1]], A[w[f], h1[f]], A[w[f], h2[f]] = b["Auh"][f]
    return A


def _column_planes(v, nl, nc, col):
    nC = nl * nc
    return np.concatenate([v[i * nC:(i + 1) * nC].reshape(nl, nc)[:, col] for i in range(4)]
                          + [v[4 * nC:].reshape(nl + 1, nc)[:, col]])


def check_implicit_diffusion_blocks(mk):
    # a column at rest in u_3 (the advection blocks then do not depend on rho e_tot)
    N, H, D0 = 12, 3000.0, 7.0
    g = G.columns(1, G.tanh_stretched_faces(N, H, 100.0))
    zc = g.zc
    rng = np.random.default_rng(4)
    rho = 1.2 * np.exp(-zc / 8000.0) * (1 + 0.01 * rng.random(N))
    T = 290.0 - 0.0065 * zc + rng.random(N)
    Y = _column_state(zc, rho, T, 3.0 + rng.standard_normal(N))
    dtg = 40.0
    o = mk(g, PAR, _num_diff(True), D0_col=np.array([D0]))
    y = Y.copy()
    o.set_precomputed(y)
    o.update_jacobian(y, dtg)
    b = o.column_blocks(0)

    def Timp(yy):
        yy = yy.copy()
        o.set_precomputed(yy)
        return o.implicit_tendency(yy)

    for plane, blk in ((3, "Aee"), (1, "Ahh")):
        for j in range(N):
            v = np.zeros_like(Y)
            v[plane * N + j] = 1.0
            eps = 1e-4 * max(abs(Y[plane * N + j]), 1.0)
            cd = dtg * (Timp(Y + eps * v) - Timp(Y - eps * v))[plane * N:(plane + 1) * N] / (2 * eps)
            cd[j] -= 1.0
            exp = np.zeros(N)
            lo, d, up = b[blk][j, 1], b[blk][j, 1], b[blk][j, 1]
            for k in range(N):  # column j of the tridiagonal block
                if k == j:
                    exp[k] = b[blk][k, 1]
                elif k == j - 1:
                    exp[k] = b[blk][k, 2]
                elif k == j + 1:
                    exp[k] = b[blk][k, 0]
            assert np.abs(cd - exp).max() <= 1e-7 * np.abs(exp).max(), (blk, j, cd, exp)


def check_implicit_diffusion_solve(mk):
    rng = np.random.default_rng(6)
    for kind in ("column", "sphere"):
        if kind == "column":
            g = G.columns(3, G.tanh_stretched_faces(16, 30000.0, 200.0))
            Y = np.concatenate([_column_planes_state(g, rng)])
            col = 1
        else:
            g = G.cubed_sphere(2, G.uniform_faces(10, 30000.0), PAR.planet_radius, PAR.Omega)
            Y = M.Smooth(PAR).state(g)
            col = 37
        nl, nc = g.nlev, g.ncol
        D0 = rng.uniform(10.0, 50.0, nc)
        dtg = 20.0
        R = rng.standard_normal(len(Y)) * np.maximum(np.abs(Y), 1.0) * 1e-3
        errs = []
        for iters in (1, 2, 4, 40):
            o = mk(g, PAR, _num_diff(True, iters), D0_col=D0)
            y = Y.copy()
            o.set_precomputed(y)
            o.update_jacobian(y, dtg)
            if iters == 1:
                A = _dense_column(o.column_blocks(col), nl)
                exact = np.linalg.solve(A, _column_planes(R, nl, nc, col))
            got = _column_planes(o.ldiv(R), nl, nc, col)
            errs.append(np.linalg.norm(got - exact) / np.linalg.norm(exact))
        assert errs[-1] < 1e-10 and errs[0] > 1e-8 and errs[1] < errs[0] and errs[2] < errs[1], (kind, errs)
        # without diffusion one iteration is the exact block arrowhead solve
        o = mk(g, PAR, P.Numerics(dt=100.0, hyperdiff=False), D0_col=D0)
        y = Y.copy()
        o.set_precomputed(y)
        o.update_jacobian(y, dtg)
        b = o.column_blocks(col)
        b["Aee"] = b["Ahh"] = np.tile([0.0, -1.0, 0.0], (nl, 1))  # the fallback -I diagonals
        A = _dense_column(b, nl)
        exact = np.linalg.solve(A, _column_planes(R, nl, nc, col))
        got = _column_planes(o.ldiv(R), nl, nc, col)
        assert np.linalg.norm(got - exact) <= 1e-11 * np.linalg.norm(exact), kind


def _column_planes_state(g, rng):
    from cadycore import initial as I

    Y = I.column_state(g, PAR, seed=3)
    nl, nc = g.nlev, g.ncol
    # w ~ 0.5 m/s on the faces (u_3 is covariant: w times the face spacing)
    jvf = np.concatenate([[g.z_face[1] - g.z_face[0]], np.diff(g.zc), [g.z_face[-1] - g.z_face[-2]]])
    Y[4 * nl * nc:] = (rng.normal(0.0, 0.5, (nl + 1, nc)) * jvf[:, None]).reshape(-1)
    return Y


def test_implicit_diffusion_moves_the_tendency():
    check_implicit_diffusion_moves_tendency(default_factory)


def test_implicit_diffusion_blocks_match_central_differences():
    check_implicit_diffusion_blocks(default_factory)


def test_approximate_arrowhead_solve_converges_to_the_dense_solve():
    check_implicit_diffusion_solve(default_factory)


ALL_CHECKS = {
    "mms_horizontal": check_mms_horizontal,
    "mms_vertical": check_mms_vertical,
    "mms_with_w": check_mms_with_w,
    "hyperdiffusion": check_hyperdiffusion,
    "vdiff_column": check_vertical_diffusion_column,
    "vdiff_sphere": check_vertical_diffusion_sphere,
    "vdiff_closures": check_vertical_diffusion_closures,
    "rayleigh_uh": check_rayleigh_uh,
    "sfc_column": check_surface_flux_column,
    "sfc_sphere": check_surface_flux_sphere,
    "idiff_move": check_implicit_diffusion_moves_tendency,
    "idiff_blocks": check_implicit_diffusion_blocks,
    "idiff_solve": check_implicit_diffusion_solve,
}
