"""Pins the CPU restatement (the oracle) against what the mathematics and the
reference's own tests fix, so that a plausible mistake (a dropped term, a
wrong sign, a transposed operand) fails here.

Reference tests mirrored here:
  * impenetrability filter            test/prognostic_equations/advection_tests.jl:21-59
  * correction + central == upwind     test/prognostic_equations/correct_implicit_advection_tests.jl:20-68
  * Rayleigh sponge profile            test/parameterized_tendencies/sponge.jl:46-144
  * R_d ~ 287                          test/parameter_tests.jl:21
"""
import math

import numpy as np
import pytest

from cadycore import geometry as G
from cadycore import initial as I
from cadycore import params as P
from oracle import Oracle, tableau

PAR = P.Params.from_toml()


@pytest.fixture(scope="module")
def sphere():
    zf = G.tanh_stretched_faces(8, 30000.0, 500.0)
    return G.cubed_sphere(3, zf, PAR.planet_radius, PAR.Omega)


@pytest.fixture(scope="module")
def sphere_oracle(sphere):
    return Oracle(sphere, PAR, P.Numerics(dt=100.0))


def _elem_poly(grid, fn):
    """Per-element polynomial of the reference coordinates, evaluated at nodes."""
    x, _, _ = G.gll()
    out = np.empty(grid.ncol)
    for e in range(grid.nelem):
        for j in range(4):
            for i in range(4):
                out[16 * e + 4 * j + i] = fn(x[i], x[j], e)
    return out


# ------------------------------------------------------------------ basics

def test_parameters_match_reference_values():
    assert abs(PAR.R_d - 287.0) / 287.0 < 1e-2  # test/parameter_tests.jl:21
    assert abs(PAR.cp_d / PAR.R_d - 3.5) < 1e-6
    assert abs(PAR.cv_d - (PAR.cp_d - PAR.R_d)) < 1e-9


def test_gll_quadrature_and_derivative_exact_on_cubics():
    x, w, D = G.gll()
    for deg in range(6):  # GLL with 4 points integrates degree <= 5 exactly
        exact = (1 - (-1) ** (deg + 1)) / (deg + 1)
        assert abs(w @ x ** deg - exact) < 1e-14
    for deg in range(4):
        np.testing.assert_allclose(D @ x ** deg, deg * x ** max(deg - 1, 0) if deg else 0 * x, atol=1e-13)


def test_tanh_stretching_bottom_layer_and_top():
    zf = G.tanh_stretched_faces(63, 30000.0, 30.0)
    assert abs(zf[1] - 30.0) < 1e-9 and zf[0] == 0.0 and zf[-1] == 30000.0
    assert np.all(np.diff(zf) > 0) and np.all(np.diff(np.diff(zf)) > 0)


def test_cubed_sphere_area_and_dss_groups(sphere):
    _, w, _ = G.gll()
    W = np.tile(np.outer(w, w).reshape(-1), sphere.nelem)
    area = (W * sphere.J_h).sum()
    assert abs(area / (4 * math.pi * PAR.planet_radius ** 2) - 1) < 1e-5
    ptr, wts = sphere.dss_ptr, sphere.dss_w
    sizes = np.diff(ptr)
    assert set(sizes.tolist()) <= {2, 3, 4}
    assert (sizes == 3).sum() == 8  # the eight cube corners
    for n in range(len(ptr) - 1):
        assert abs(wts[ptr[n]:ptr[n + 1]].sum() - 1) < 1e-14
    # unique node count of a cubed sphere: 6 h^2 (nq-1)^2 + 2
    assert sphere.n_nodes == 6 * 9 * 9 + 2


# ------------------------------------------------- spectral element operators

def test_strong_gradient_exact_on_element_polynomials(sphere, sphere_oracle):
    f = _elem_poly(sphere, lambda a, b, e: a ** 3 + 2 * a * b ** 2 - b + e)
    g1, g2 = sphere_oracle.op("grad", 0, f)
    np.testing.assert_allclose(g1, _elem_poly(sphere, lambda a, b, e: 3 * a ** 2 + 2 * b ** 2), atol=1e-12)
    np.testing.assert_allclose(g2, _elem_poly(sphere, lambda a, b, e: 4 * a * b - 1), atol=1e-12)


def test_curl_of_gradient_and_div_of_curl_vanish(sphere, sphere_oracle):
    rng = np.random.default_rng(0)
    f = rng.standard_normal(sphere.ncol)
    g1, g2 = sphere_oracle.op("grad", 2, f)
    c3 = sphere_oracle.op("curl3", 2, g1, g2)
    assert np.abs(c3).max() * np.abs(sphere.J_h).max() < 1e-10 * np.abs(g1).max()
    a1, a2 = sphere_oracle.op("curl12", 2, f)
    d = sphere_oracle.op("div", 2, a1, a2)
    assert np.abs(d).max() < 1e-12 * np.abs(a1).max()


def _IP(grid, k, a, b):
    """Discrete inner product sum WJ a b at centre level k (3-D Jacobian)."""
    _, w, _ = G.gll()
    zc, Jvc, _ = G.vertical_metrics(grid.z_face)
    s = grid.scale(zc[k])
    W = np.tile(np.outer(w, w).reshape(-1), grid.nelem)
    return np.sum(W * grid.J_h * s * s * Jvc[k] * a * b)


def _contra(grid, k, u1, u2):
    zc, _, _ = G.vertical_metrics(grid.z_face)
    s2 = grid.scale(zc[k]) ** 2
    g = grid.gup
    return (g[0] * u1 + g[1] * u2) / s2, (g[1] * u1 + g[2] * u2) / s2


def test_weak_divergence_is_negative_adjoint_of_strong_gradient(sphere, sphere_oracle):
    rng = np.random.default_rng(1)
    k = 3
    phi = rng.standard_normal(sphere.ncol)
    v1, v2 = rng.standard_normal(sphere.ncol), rng.standard_normal(sphere.ncol)
    wd = sphere_oracle.op("wdiv", k, v1, v2)
    g1, g2 = sphere_oracle.op("grad", k, phi)
    lhs = _IP(sphere, k, phi, wd)
    rhs = -_IP(sphere, k, v1, g1) - _IP(sphere, k, v2, g2)
    assert abs(lhs - rhs) < 1e-10 * abs(rhs)


def test_weak_gradient_is_negative_adjoint_of_strong_divergence(sphere, sphere_oracle):
    rng = np.random.default_rng(2)
    k = 1
    psi = rng.standard_normal(sphere.ncol)
    v1, v2 = rng.standard_normal(sphere.ncol), rng.standard_normal(sphere.ncol)
    d = sphere_oracle.op("div", k, v1, v2)
    w1, w2 = sphere_oracle.op("wgrad", k, psi)
    lhs = _IP(sphere, k, psi, d)
    rhs = -_IP(sphere, k, v1, w1) - _IP(sphere, k, v2, w2)
    assert abs(lhs - rhs) < 1e-10 * abs(rhs)


def test_weak_curls_are_adjoints_of_strong_curls(sphere, sphere_oracle):
    rng = np.random.default_rng(3)
    k = 2
    phi1, phi2 = rng.standard_normal(sphere.ncol), rng.standard_normal(sphere.ncol)
    psi = rng.standard_normal(sphere.ncol)
    # <phi, wcurl12(psi)>  (C12 paired with CT12)  ==  <curl3(phi), psi>  (CT3 with C3 psi g_33)
    w1, w2 = sphere_oracle.op("wcurl12", k, psi)
    c3 = sphere_oracle.op("curl3", k, phi1, phi2)
    lhs = _IP(sphere, k, phi1, w1) + _IP(sphere, k, phi2, w2)
    rhs = _IP(sphere, k, c3, psi)
    assert abs(lhs - rhs) < 1e-10 * abs(rhs)
    wc3 = sphere_oracle.op("wcurl3", k, phi1, phi2)
    a1, a2 = sphere_oracle.op("curl12", k, psi)
    lhs = _IP(sphere, k, psi, wc3)
    rhs = _IP(sphere, k, phi1, a1) + _IP(sphere, k, phi2, a2)
    assert abs(lhs - rhs) < 1e-10 * abs(rhs)


def test_weak_operators_consistent_at_interior_nodes(sphere, sphere_oracle):
    f = _elem_poly(sphere, lambda a, b, e: a ** 2 * b - 3 * b ** 3 + a)
    w1, w2 = sphere_oracle.op("wgrad", 0, f)
    g1, g2 = sphere_oracle.op("grad", 0, f)
    interior = np.tile(np.array([0, 0, 0, 0, 0, 1, 1, 0, 0, 1, 1, 0, 0, 0, 0, 0], bool), sphere.nelem)
    np.testing.assert_allclose(w1[interior], g1[interior], atol=1e-11)
    np.testing.assert_allclose(w2[interior], g2[interior], atol=1e-11)


def test_global_weak_divergence_integrates_to_zero_after_dss(sphere, sphere_oracle):
    """Discrete divergence theorem on the closed sphere: sum WJ wdiv(F) = 0
    for a continuous flux, with the DSS between the two."""
    g = sphere
    X = g.basis  # build a continuous tangent field from a global vector
    cart = np.stack([np.cos(g.lat) * np.cos(g.lon), np.cos(g.lat) * np.sin(g.lon), np.sin(g.lat)])
    Vg = np.cross(np.array([0.3, -1.0, 0.5]), cart.T).T + 0.2 * cart[2] * np.cross(cart.T, [1, 0, 0]).T
    u1 = np.einsum("ic,ic->c", Vg, X[6:9])
    u2 = np.einsum("ic,ic->c", Vg, X[9:12])
    k = 0
    v1, v2 = _contra(g, k, u1, u2)
    wd = sphere_oracle.op("wdiv", k, v1, v2)
    total = _IP(g, k, np.ones(g.ncol), wd)
    scale = _IP(g, k, np.ones(g.ncol), np.abs(wd))
    assert abs(total) < 1e-10 * scale


def test_dss_idempotent_and_identity_on_continuous_fields(sphere, sphere_oracle):
    o = sphere_oracle
    g = sphere
    rng = np.random.default_rng(4)
    Y = rng.standard_normal(o.S)
    o.dss(Y)
    Y2 = Y.copy()
    o.dss(Y2)
    np.testing.assert_allclose(Y2, Y, rtol=0, atol=1e-12)
    # a continuous scalar and a continuous tangent vector are left unchanged
    nl = g.nlev
    cart = np.stack([np.cos(g.lat) * np.cos(g.lon), np.cos(g.lat) * np.sin(g.lon), np.sin(g.lat)])
    scal = np.tile(np.sin(3 * g.lat) * np.cos(g.lon), (nl, 1))
    V = np.cross(np.array([0.0, 0.0, 1.0]), cart.T).T
    u1 = np.tile(np.einsum("ic,ic->c", V, g.basis[6:9]), (nl, 1))
    u2 = np.tile(np.einsum("ic,ic->c", V, g.basis[9:12]), (nl, 1))
    Yc = I.pack(scal, u1, u2, 2 * scal, np.zeros((nl + 1, g.ncol)))
    Yd = Yc.copy()
    o.dss(Yd)
    np.testing.assert_allclose(Yd, Yc, rtol=0, atol=1e-9 * np.abs(Yc).max())


# ------------------------------------------------------------ column physics

def test_impenetrability_filter_zeroes_boundary_u3(sphere, sphere_oracle):
    o = sphere_oracle
    Y = I.sphere_state(sphere, PAR, seed=5)
    rho, u1, u2, rhoe, u3 = I.unpack(Y, sphere.ncol, sphere.nlev)
    u3[:] = np.random.default_rng(5).standard_normal(u3.shape)
    o.set_precomputed(Y)
    u3b = I.unpack(Y, sphere.ncol, sphere.nlev)[4]
    assert np.all(u3b[0] == 0) and np.all(u3b[-1] == 0)
    assert np.array_equal(u3b[1:-1], u3[1:-1])


def test_dry_equation_of_state_and_pressure_derivative():
    """dp/drho at fixed rho e_tot and velocity is kappa (T_0 cp_d - K - Phi)
    (manual_sparse_jacobian.jl:816-818): a finite difference of the oracle's
    precomputed pressure must reproduce it, which pins e_int = cv_d (T - T_0)
    - R_d T_0 and p = rho R_d T."""
    zf = G.tanh_stretched_faces(10, 30000.0, 500.0)
    g = G.columns(1, zf)
    o = Oracle(g, PAR, P.Numerics(hyperdiff=False, dt=10.0))
    Y = I.column_state(g, PAR, seed=0)
    o.set_precomputed(Y)
    K, T, p, h = o.precomputed()
    rho, rhoe = Y[:10], Y[30:40]
    e_int = rhoe / rho - K - PAR.grav * g.zc
    np.testing.assert_allclose(e_int, PAR.cv_d * (T - PAR.T_0) - PAR.R_d * PAR.T_0, rtol=1e-12)
    np.testing.assert_allclose(p, rho * PAR.R_d * T, rtol=1e-13)
    np.testing.assert_allclose(h, rhoe / rho + PAR.R_d * T, rtol=1e-13)
    kap = PAR.R_d / PAR.cv_d
    eps = 1e-6
    Yp = Y.copy()
    Yp[:10] *= 1 + eps
    o.set_precomputed(Yp)
    _, _, pp, _ = o.precomputed()
    Ym = Y.copy()
    Ym[:10] *= 1 - eps
    o.set_precomputed(Ym)
    _, _, pm, _ = o.precomputed()
    # K depends on rho only through nothing (velocities fixed): exact derivative
    dpdr = (pp - pm) / (2 * eps * rho)
    np.testing.assert_allclose(dpdr, kap * (PAR.T_0 * PAR.cp_d - K - PAR.grav * g.zc), rtol=1e-7)


def _reference_state_column(nl=20, ztop=30000.0):
    """Discrete state with T = T_r(Pi), Phi - Phi_r(Pi) = 0 level by level and
    no wind: the implicit u_3 tendency must vanish identically because
    grad(Phi_r) + cp_d theta_vr grad(Pi) == 0 (refstate_thermodynamics.jl)."""
    zf = G.tanh_stretched_faces(nl, ztop, 200.0)
    g = G.columns(1, zf)
    zc = g.zc
    cp, R = PAR.cp_d, PAR.R_d
    Tm, Ts = PAR.T_min_ref, PAR.T_surf_ref

    def phir(Pi):
        return -cp * (Tm * math.log(Pi) + (Ts - Tm) / 7 * (Pi ** 7 - 1))

    Pis = []
    for z in zc:
        lo, hi = 1e-6, 1.0
        for _ in range(200):
            mid = 0.5 * (lo + hi)
            if phir(mid) > PAR.grav * z:
                lo = mid
            else:
                hi = mid
        Pis.append(0.5 * (lo + hi))
    Pi = np.array(Pis)
    T = Tm + (Ts - Tm) * Pi ** 7
    p = PAR.p_ref_theta * Pi ** (cp / R)
    rho = p / (R * T)
    rhoe = rho * (PAR.cv_d * (T - PAR.T_0) - R * PAR.T_0 + PAR.grav * zc)
    Y = I.pack(rho[:, None], 0 * rho[:, None], 0 * rho[:, None], rhoe[:, None], np.zeros((nl + 1, 1)))
    return g, Y


def test_reference_state_is_exactly_balanced_in_implicit_tendency():
    g, Y = _reference_state_column()
    o = Oracle(g, PAR, P.Numerics(hyperdiff=False, dt=10.0))
    o.set_precomputed(Y)
    Yt = o.implicit_tendency(Y)
    u3t = I.unpack(Yt, 1, g.nlev)[4]
    # compare with the size of either term (g dz ~ 10 m/s^2 per level step)
    assert np.abs(u3t).max() < 1e-9 * PAR.grav * np.diff(g.zc).max()
    # and a perturbation of T breaks it with the right sign (warm air rises)
    rho, _, _, rhoe, _ = I.unpack(Y, 1, g.nlev)
    Y2 = Y.copy()
    r2 = I.unpack(Y2, 1, g.nlev)
    r2[3][5, 0] += rho[5, 0] * PAR.cv_d * 1.0  # +1 K at level 5 at fixed rho
    o.set_precomputed(Y2)
    u3t2 = I.unpack(o.implicit_tendency(Y2), 1, g.nlev)[4]
    assert u3t2[5, 0] > 0 or u3t2[6, 0] > 0


def test_mass_conserved_by_full_step(sphere, sphere_oracle):
    o = sphere_oracle
    Y = I.sphere_state(sphere, PAR, seed=1)
    o.dss(Y)
    _, w, _ = G.gll()
    zc, Jvc, _ = G.vertical_metrics(sphere.z_face)
    s = sphere.scale(zc)
    W = np.tile(np.outer(w, w).reshape(-1), sphere.nelem)
    J = (W * sphere.J_h)[None, :] * (s ** 2 * Jvc)[:, None]
    m0 = (I.unpack(Y, sphere.ncol, sphere.nlev)[0] * J).sum()
    o.step(Y, 100.0, 2)
    m1 = (I.unpack(Y, sphere.ncol, sphere.nlev)[0] * J).sum()
    assert abs(m1 / m0 - 1) < 1e-13


def test_column_energy_conserved_by_vertical_transport():
    """rho e_tot changes only through boundary fluxes, which are zero: the
    column integral of rho e_tot J is conserved by the implicit transport."""
    zf = G.tanh_stretched_faces(12, 30000.0, 500.0)
    g = G.columns(2, zf)
    o = Oracle(g, PAR, P.Numerics(hyperdiff=False, dt=10.0))
    Y = I.column_state(g, PAR, seed=3)
    r = I.unpack(Y, 2, 12)
    r[4][1:-1] = np.random.default_rng(3).standard_normal((11, 2)) * 50
    o.set_precomputed(Y)
    Yt = o.implicit_tendency(Y)
    zc, Jvc, _ = G.vertical_metrics(zf)
    rt = I.unpack(Yt, 2, 12)
    for comp in (0, 3):
        tot = (rt[comp] * Jvc[:, None]).sum(axis=0)
        ref = (np.abs(rt[comp]) * Jvc[:, None]).sum(axis=0)
        assert np.all(np.abs(tot) < 1e-12 * ref)


# --------------------------------------------------- Jacobian and the solve

def _column_setup(nl=12, seed=7, rayleigh=False):
    zf = G.tanh_stretched_faces(nl, 30000.0, 400.0)
    g = G.columns(1, zf)
    n = P.Numerics(hyperdiff=False, dt=30.0, rayleigh=rayleigh)
    o = Oracle(g, PAR, n)
    Y = I.column_state(g, PAR, seed=seed)
    r = I.unpack(Y, 1, nl)
    r[4][1:-1, 0] = np.random.default_rng(seed).standard_normal(nl - 1) * 20.0
    return g, o, Y


@pytest.mark.parametrize("rayleigh", [False, True])
def test_jacobian_exact_blocks_match_central_differences(rayleigh):
    """The blocks that the reference derives exactly (u3 rows and the rho row)
    match a central-difference column Jacobian of the oracle's own implicit
    tendency -- the reference's debugging method (implicit/jacobian.jl:98-155)."""
    nl = 12
    g, o, Y = _column_setup(nl, rayleigh=rayleigh)
    o.set_precomputed(Y)
    dtg = 7.0
    o.update_jacobian(Y, dtg)
    B = o.column_blocks(0)
    nC = nl
    off = {"rho": 0, "u1": nC, "u2": 2 * nC, "rhoe": 3 * nC, "u3": 4 * nC}

    def Timp(Yx):
        Yc = Yx.copy()
        o.set_precomputed(Yc)
        return o.implicit_tendency(Yc)

    def col_fd(comp, idx, h):
        Yp, Ym = Y.copy(), Y.copy()
        Yp[off[comp] + idx] += h
        Ym[off[comp] + idx] -= h
        return (Timp(Yp) - Timp(Ym)) / (2 * h)

    rho = Y[:nC]

    def Tac(Yx, comp, idx, h):
        """acoustic face tendency -grad_v(p) / I^f(rho) of a perturbed state"""
        Yc = Yx.copy()
        Yc[off[comp] + idx] += h
        o.set_precomputed(Yc)
        _, _, p, _ = o.precomputed()
        r = Yc[:nC]
        out = np.zeros(nl + 1)
        out[1:nl] = -(p[1:] - p[:-1]) / (0.5 * (r[1:] + r[:-1]))
        return out

    # d(T_rho)/d(u3_f): A(rho,u3)/dtg ; d(T_u3)/d(rho_c), d(T_u3)/d(rhoe_c), d(T_u3)/d(u3)
    for f in range(1, nl):
        d = col_fd("u3", f, 1e-3)
        for kk, jj in ((f - 1, 1), (f, 0)):
            np.testing.assert_allclose(d[off["rho"] + kk], B["Aru"][kk, jj] / dtg, rtol=1e-6, atol=1e-14)
        # (u3, u3): the reference keeps only the acoustic path through K
        # (manual_sparse_jacobian.jl:857-868); it equals the derivative of
        # -grad_v(p) / I^f(rho) with respect to u3, taken here by differences
        # of the oracle's own pressure.
        # K is quadratic in u3 and p linear in K: central differences are exact
        # for any step, so a large one keeps the pressure difference well
        # above round-off
        dac = (Tac(Y, "u3", f, 1e3) - Tac(Y, "u3", f, -1e3)) / 2e3
        beta = 0.0
        if rayleigh:
            zd, zmax = PAR.zd_rayleigh, g.z_face[-1]
            zf_ = g.z_face[f]
            beta = PAR.alpha_rayleigh_w * np.sin(np.pi * (zf_ - zd) / (2 * (zmax - zd))) ** 2 if zf_ > zd else 0.0
        np.testing.assert_allclose(dac[f], (B["di"][f] + 1) / dtg + beta, rtol=1e-5, atol=1e-15)
        np.testing.assert_allclose(dac[f - 1], B["up"][f - 1] / dtg if f - 1 >= 1 else 0.0, rtol=1e-5, atol=1e-15)
        if f + 1 < nl:
            np.testing.assert_allclose(dac[f + 1], B["lo"][f + 1] / dtg, rtol=1e-5, atol=1e-15)
    for k in range(nl):
        # the acoustic path alone is linear in rho e_tot: exact
        he = 1e-3 * Y[off["rhoe"] + k]
        dac = (Tac(Y, "rhoe", k, he) - Tac(Y, "rhoe", k, -he)) / (2 * he)
        for f, jj in ((k, 1), (k + 1, 0)):
            if 1 <= f <= nl - 1:
                np.testing.assert_allclose(dac[f], B["Aue"][f, jj] / dtg, rtol=1e-7)


def test_jacobian_buoyancy_blocks_converge_to_discrete_derivative():
    """The (u3, rho) and (u3, rho e_tot) blocks carry the continuum
    linearisation of the buoyancy (manual_sparse_jacobian.jl:790-824); against
    central differences of the discrete tendency they must agree to the
    truncation error, shrinking at second order as the column is refined
    (a dropped term or a sign error stays O(1))."""
    errs = {}
    for nl in (16, 32, 64):
        zf = G.uniform_faces(nl, 30000.0)
        g = G.columns(1, zf)
        o = Oracle(g, PAR, P.Numerics(hyperdiff=False, dt=30.0))
        Y = I.column_state(g, PAR, seed=7)
        r = I.unpack(Y, 1, nl)
        zc = g.zc
        r[4][1:-1, 0] = 20.0 * np.sin(2 * np.pi * g.z_face[1:-1] / 30000.0) * np.diff(zc) / 1000.0
        o.set_precomputed(Y)
        dtg = 7.0
        o.update_jacobian(Y, dtg)
        B = o.column_blocks(0)

        def Timp(Yx):
            Yc = Yx.copy()
            o.set_precomputed(Yc)
            return o.implicit_tendency(Yc)

        worst = {"Aur": 0.0, "Aue": 0.0}
        scale = {"Aur": np.abs(B["Aur"]).max() / dtg, "Aue": np.abs(B["Aue"]).max() / dtg}
        for k in range(nl):
            for comp, blk in ((0, "Aur"), (3 * nl, "Aue")):
                h = 1e-7 * abs(Y[comp + k])
                Yp, Ym = Y.copy(), Y.copy()
                Yp[comp + k] += h
                Ym[comp + k] -= h
                d = (Timp(Yp) - Timp(Ym)) / (2 * h)
                for f, jj in ((k, 1), (k + 1, 0)):
                    if 1 <= f <= nl - 1:
                        worst[blk] = max(worst[blk], abs(d[4 * nl + f] - B[blk][f, jj] / dtg) / scale[blk])
        errs[nl] = worst
    for blk in ("Aur", "Aue"):
        assert errs[64][blk] < 0.02
        assert errs[16][blk] / errs[32][blk] > 3.0 and errs[32][blk] / errs[64][blk] > 3.0


def test_schur_solve_matches_dense_block_solve():
    nl = 10
    g, o, Y = _column_setup(nl)
    o.set_precomputed(Y)
    dtg = 20.0
    o.update_jacobian(Y, dtg)
    B = o.column_blocks(0)
    nC, nF = nl, nl + 1
    N = 4 * nC + nF
    A = np.zeros((N, N))
    r_rho, r_u1, r_u2, r_e, r_u3 = 0, nC, 2 * nC, 3 * nC, 4 * nC
    for q in range(4 * nC):
        A[q, q] = -1.0
    for k in range(nl):
        for jj, f in ((0, k), (1, k + 1)):
            A[r_rho + k, r_u3 + f] += B["Aru"][k, jj]
            A[r_e + k, r_u3 + f] += B["Aeu"][k, jj]
    for f in range(nF):
        A[r_u3 + f, r_u3 + f] += B["di"][f]
        if f > 0:
            A[r_u3 + f, r_u3 + f - 1] += B["lo"][f]
            A[r_u3 + f, r_rho + f - 1] += B["Aur"][f, 0]
            A[r_u3 + f, r_e + f - 1] += B["Aue"][f, 0]
        if f < nl:
            A[r_u3 + f, r_u3 + f + 1] += B["up"][f]
            A[r_u3 + f, r_rho + f] += B["Aur"][f, 1]
            A[r_u3 + f, r_e + f] += B["Aue"][f, 1]
    R = np.random.default_rng(9).standard_normal(N)
    R[r_u1:r_e] = 0.0  # no u_h residual in the first Newton iteration
    dY = o.ldiv(R)
    dense = np.linalg.solve(A, R)
    np.testing.assert_allclose(dY, dense, rtol=0, atol=1e-9 * np.abs(dense).max())


def test_ars343_tableau_order_conditions_and_convergence():
    aE, aI, b, c = tableau()
    np.testing.assert_allclose(aE.sum(1), c, atol=1e-14)
    np.testing.assert_allclose(aI.sum(1), c, atol=1e-14)
    assert abs(b.sum() - 1) < 1e-14
    assert abs(b @ c - 0.5) < 1e-14
    assert abs(b @ c ** 2 - 1 / 3) < 1e-14
    assert abs(b @ aE @ c - 1 / 6) < 1e-13
    assert abs(b @ aI @ c - 1 / 6) < 1e-13
    g = aI[1, 1]
    assert abs(6 * g ** 3 - 18 * g ** 2 + 9 * g - 1) < 1e-14
    # ARS (3,4,3) published decimals (Ascher, Ruuth & Spiteri 1997)
    for v, ref in ((aE[2, 0], 0.3212788860), (aE[2, 1], 0.3966543747), (aE[3, 0], -0.105858296),
                   (aE[3, 1], 0.5529291479), (b[1], 1.208496649), (b[2], -0.644363171)):
        assert abs(v - ref) < 2e-9
    # third-order convergence of the stage structure used by the step on
    # y' = lam_E y + lam_I y (implicit tendency back-computed as (U - U~)/dtg)
    lamE, lamI = 0.7 + 0.3j, -2.0

    def run(dt, n):
        y = 1.0 + 0j
        for _ in range(n):
            FE, FI = [0] * 4, [0] * 4
            FE[0] = lamE * y
            for i in range(1, 4):
                ut = y + dt * sum(aE[i, j] * FE[j] + (aI[i, j] * FI[j] if j else 0) for j in range(i))
                u = ut / (1 - dt * aI[i, i] * lamI)
                FI[i] = (u - ut) / (dt * aI[i, i])
                FE[i] = lamE * u
            y = y + dt * sum(b[j] * (FE[j] + (FI[j] if j else 0)) for j in range(4))
        return y

    exact = np.exp((lamE + lamI) * 1.0)
    e1 = abs(run(0.05, 20) - exact)
    e2 = abs(run(0.025, 40) - exact)
    assert 2.8 < math.log2(e1 / e2) < 3.3


def test_ars343_final_combination_from_the_last_stage():
    """ARS343's implicit tableau is stiffly accurate (b = last row of a_I), so
    with F_I back-computed as (U - U~) / (dt a_I[i][i]) the step's final
    combination u + dt sum b_j (F_E,j + F_I,j) equals U_4 + dt sum_j (b_j -
    a_E[3][j]) F_E,j, with a_E[3][3] = 0: the form in which the native step
    forms it inside its last explicit tendency.  Checked on a nonlinear
    vector system (not only y' = lam y), and by a plausible slip failing it
    (dropping the a_E[3][j] terms)."""
    aE, aI, b, c = tableau()
    np.testing.assert_array_equal(aI[3], b)
    assert aE[3, 3] == 0.0 and aI[0].sum() == 0.0
    rng = np.random.default_rng(5)
    M = rng.standard_normal((6, 6)) * 0.3
    lin = -np.diag(rng.uniform(0.5, 3.0, 6))

    def fe(y):  # explicit tendency: nonlinear
        return np.tanh(M @ y) + 0.1 * y ** 2

    dt, u = 0.2, rng.standard_normal(6)
    FE, FI, U = [None] * 4, [None] * 4, [u]
    FE[0] = fe(u)
    for i in range(1, 4):
        ut = u + dt * sum(aE[i, j] * FE[j] for j in range(i)) + dt * sum(aI[i, j] * FI[j] for j in range(1, i))
        Ui = np.linalg.solve(np.eye(6) - dt * aI[i, i] * lin, ut)  # the implicit solve
        FI[i] = (Ui - ut) / (dt * aI[i, i])
        FE[i] = fe(Ui)
        U.append(Ui)
    ref = u + dt * sum(b[j] * (FE[j] + (FI[j] if j else 0)) for j in range(4))
    fused = U[3] + dt * sum((b[j] - aE[3, j]) * FE[j] for j in range(4))
    np.testing.assert_allclose(fused, ref, rtol=1e-13, atol=1e-13)
    slip = U[3] + dt * sum(b[j] * FE[j] for j in range(4))
    assert np.abs(slip - ref).max() > 1e-3


def test_rayleigh_sponge_profile_closed_form():
    """beta = alpha sin^2(pi (z - zd) / (2 (zmax - zd))) above zd
    (rayleigh_sponge.jl:7-28, sponge.jl test)."""
    nl = 12
    zf = G.uniform_faces(nl, 30000.0)
    g = G.columns(1, zf)
    o = Oracle(g, PAR, P.Numerics(hyperdiff=False, dt=30.0, rayleigh=True))
    Y = I.column_state(g, PAR, seed=1)
    r = I.unpack(Y, 1, nl)
    r[4][1:-1, 0] = 1.0
    o.set_precomputed(Y)
    t_on = I.unpack(o.implicit_tendency(Y), 1, nl)[4][:, 0]
    o2 = Oracle(g, PAR, P.Numerics(hyperdiff=False, dt=30.0, rayleigh=False))
    Y2 = Y.copy()
    o2.set_precomputed(Y2)
    t_off = I.unpack(o2.implicit_tendency(Y2), 1, nl)[4][:, 0]
    zd, zmax = PAR.zd_rayleigh, 30000.0
    beta = np.where(zf > zd, PAR.alpha_rayleigh_w * np.sin(np.pi * (zf - zd) / (2 * (zmax - zd))) ** 2, 0.0)
    np.testing.assert_allclose((t_on - t_off)[1:-1], -beta[1:-1] * 1.0, atol=1e-12)


# ------------------------------------------------------------------ upwinding

def test_correction_plus_central_equals_upwind_and_vanleer_is_bounded():
    """correct_implicit_advection_tests.jl:20-68: correction + central ==
    upwind; van Leer face values stay within the neighbouring cell values."""
    nl = 16
    zf = G.uniform_faces(nl, 16000.0)
    g = G.columns(1, zf)
    for scheme in (P.UPWIND_FIRST, P.UPWIND_VANLEER):
        o = Oracle(g, PAR, P.Numerics(hyperdiff=False, dt=5.0, upwinding=scheme))
        Y = I.column_state(g, PAR, seed=2)
        r = I.unpack(Y, 1, nl)
        rng = np.random.default_rng(11)
        r[4][1:-1, 0] = rng.standard_normal(nl - 1) * 300.0
        o.set_precomputed(Y)
        corr = o.correct_implicit(Y)
        cen = o.implicit_tendency(Y)
        rc = I.unpack(corr, 1, nl)
        assert np.all(rc[0] == 0) and np.all(rc[4] == 0)
        # upwind transport recomputed independently from the face fluxes
        K, T, p, h = o.precomputed()
        rho = r[0][:, 0]
        zc, Jvc, Jvf = G.vertical_metrics(zf)
        u3 = r[4][:, 0] / Jvf ** 2
        F = np.zeros(nl + 1)
        for f in range(1, nl):
            rJ = 0.5 * (rho[f - 1] * Jvc[f - 1] + rho[f] * Jvc[f])
            up = h[f - 1] if u3[f] >= 0 else h[f]
            if scheme == P.UPWIND_VANLEER and 2 <= f <= nl - 2:
                # bounded by the neighbouring values of the upwind cell
                kk = f - 1 if u3[f] >= 0 else f
                val = None  # value checked via bounds below
            F[f] = rJ * u3[f] * up
        if scheme == P.UPWIND_FIRST:
            vtt_up = -(F[1:] - F[:-1]) / Jvc
            np.testing.assert_allclose(rc[3][:, 0] + I.unpack(cen, 1, nl)[3][:, 0], vtt_up, rtol=1e-10,
                                       atol=1e-12 * np.abs(vtt_up).max())
        else:
            # recover face values from the corrected transport and check bounds
            tot = rc[3][:, 0] + I.unpack(cen, 1, nl)[3][:, 0]
            flux = np.zeros(nl + 1)
            for k in range(nl):
                flux[k + 1] = flux[k] - tot[k] * Jvc[k]
            for f in range(1, nl):
                rJ = 0.5 * (rho[f - 1] * Jvc[f - 1] + rho[f] * Jvc[f])
                if abs(u3[f]) < 1e-12:
                    continue
                face = flux[f] / (rJ * u3[f])
                lo, hi = min(h[f - 1], h[f]), max(h[f - 1], h[f])
                span = hi - lo
                assert lo - 1e-9 * abs(hi) - 1e-12 * span <= face <= hi + 1e-9 * abs(hi)


def test_storage_rounding_switch():
    """Off by default (stepping twice from one state is bitwise reproducible
    and leaves values that float32 cannot hold); on, the stepped state is
    float32-representable and differs from the unrounded one at round-off
    level."""
    zf = G.tanh_stretched_faces(12, 30000.0, 500.0)
    g = G.columns(2, zf)
    num = P.Numerics(hyperdiff=False, dt=10.0)
    Y0 = I.column_state(g, PAR, seed=3)

    def run(rnd):
        o = Oracle(g, PAR, num)
        Y = Y0.copy()
        o.dss(Y)
        o.set_precomputed(Y)
        if rnd:
            o.set_storage_rounding(True)
        o.step(Y, 10.0, 3)
        return Y

    a, a2, b = run(False), run(False), run(True)
    np.testing.assert_array_equal(a, a2)
    assert np.any(a != a.astype(np.float32).astype(np.float64))
    np.testing.assert_array_equal(b, b.astype(np.float32).astype(np.float64))
    rel = np.linalg.norm(a - b) / np.linalg.norm(a)
    assert 0 < rel < 1e-5
