"""Manufactured smooth states on the deep cubed sphere and their analytic
explicit tendencies (test infrastructure for the oracle pinning tests).

Every field is an analytic function of the unit position vector ``n`` and the
height ``z``.  The continuous explicit tendency of the dry equations in the
reference's vector-invariant form (``src/prognostic_equations/advection.jl``
:36-91 horizontal dynamics, :205-290 vertical advection; continuous form in
``docs/src/equations.md``) is evaluated from those functions with central
differences of step ``1e-5`` in ``n`` (or ``1e-2`` m in ``x``), whose error
(~1e-10 relative) is far below the discretisation error being measured:

    rho_t    = -div_h(rho V)
    rhoe_t   = -div_h(rho V h_tot),  h_tot = cp_d (T - T_0) + K + Phi
    V_t,h    = -(f + zeta) n x V - grad_h K - grad_h p / rho       (w = 0)
    w_t      = -((omega + 2 Omega) x V) . n - dK/dz                (explicit part)

with ``f = 2 Omega . n``, ``zeta = n . curl V`` and ``omega = curl V`` (3-D).
The state is given to the oracle in covariant components: ``u_a = s V . a_a(R)``
with ``s = (R + z) / R`` and ``u_3 = Jv w``; tendencies come back the same way.
"""
from __future__ import annotations

import numpy as np

DN = 1e-5  # central-difference step in the unit vector n


def unit_vectors(grid):
    lat, lon = grid.lat, grid.lon
    return np.stack([np.cos(lat) * np.cos(lon), np.cos(lat) * np.sin(lon), np.sin(lat)])


def proj(n, v):
    """Tangential part of v at unit vectors n (both (3, N))."""
    return v - n * np.einsum("in,in->n", n, v)[None, :] / np.einsum("in,in->n", n, n)[None, :]


def grad_n(f, n):
    """Ambient gradient of the scalar function f(n) -> (N,), shape (3, N)."""
    out = np.empty_like(n)
    for k in range(3):
        e = np.zeros((3, 1))
        e[k] = DN
        out[k] = (f(n + e) - f(n - e)) / (2 * DN)
    return out


def jac_n(F, n):
    """Ambient Jacobian dF_i/dn_k of a vector function F(n) -> (3, N); (3, 3, N)."""
    out = np.empty((3, 3, n.shape[1]))
    for k in range(3):
        e = np.zeros((3, 1))
        e[k] = DN
        out[:, k] = (F(n + e) - F(n - e)) / (2 * DN)
    return out


def surf_grad(f, n, r):
    """Surface gradient on the sphere of radius r."""
    return proj(n, grad_n(f, n)) / r


def surf_div(F, n, r):
    """Surface divergence tr(P dF/dx) of a tangent field on the sphere of radius r."""
    J = jac_n(F, n)
    P = np.eye(3)[:, :, None] - n[:, None, :] * n[None, :, :]
    return np.einsum("ikn,kin->n", P, J) / r


def cross(a, b):
    return np.cross(a.T, b.T).T


class Smooth:
    """A smooth non-axisymmetric dry state: rho, T, V (tangent) as functions of
    (n, z).  ``w_amp`` adds a vertical velocity w(n, z)."""

    def __init__(self, params, U0=25.0, rho_amp=0.05, T_amp=6.0, w_amp=0.0):
        self.p = params
        self.U0, self.ra, self.Ta, self.wa = U0, rho_amp, T_amp, w_amp

    # horizontal shapes (polynomials of n, evaluated for any n)
    @staticmethod
    def _Prho(n):
        return n[0] * n[1] + 0.5 * n[2] ** 2 - 0.3 * n[0]

    @staticmethod
    def _PT(n):
        return n[0] * n[2] + 0.4 * n[1] - 0.2 * n[1] * n[2]

    # stream function and velocity potential: polynomials with their gradients
    # written out, so that a tendency needs one level of differencing only
    @staticmethod
    def _grad_psi(n):  # psi = n_z + 0.3 n_x n_y + 0.2 n_y n_z^2
        return np.stack([0.3 * n[1], 0.3 * n[0] + 0.2 * n[2] ** 2, 1.0 + 0.4 * n[1] * n[2]])

    @staticmethod
    def _grad_chi(n):  # chi = 0.4 n_x n_z - 0.25 n_y + 0.1 n_x n_y n_z
        return np.stack([0.4 * n[2] + 0.1 * n[1] * n[2], -0.25 + 0.1 * n[0] * n[2],
                         0.4 * n[0] + 0.1 * n[0] * n[1]])

    def rho(self, n, z):
        return 1.2 * np.exp(-z / 8000.0) * (1.0 + self.ra * self._Prho(n))

    def T(self, n, z):
        return 285.0 - 0.004 * z + self.Ta * self._PT(n) * (1.0 + z / 30000.0)

    def Uz(self, z):
        return self.U0 * (1.0 + z / 20000.0)

    def V(self, n, z):
        """Tangent velocity (3, N) at unit vectors n and height z (scalar or (N,))."""
        nn = n / np.sqrt(np.einsum("in,in->n", n, n))[None, :]
        rot = cross(nn, self._grad_psi(n))
        div = proj(nn, self._grad_chi(n))
        return self.Uz(z) * (rot + 0.5 * div)

    def w(self, n, z):
        return self.wa * (0.3 + n[0] * n[1] - 0.5 * n[2]) * np.sin(np.pi * z / 30000.0)

    def p_of(self, n, z):
        return self.rho(n, z) * self.p.R_d * self.T(n, z)

    def K(self, n, z):
        V = self.V(n, z)
        return 0.5 * (np.einsum("in,in->n", V, V) + self.w(n, z) ** 2)

    def h_tot(self, n, z):
        P = self.p
        return P.cp_d * (self.T(n, z) - P.T_0) + self.K(n, z) + P.grav * z

    # ---------------------------------------------------------- the state
    def state(self, grid):
        """Native flat state [rho | u1 | u2 | rhoe | u3] (float64)."""
        P = self.p
        n = unit_vectors(grid)
        zc = grid.zc
        zf = grid.z_face
        _, _, Jvf = grid.vertical()
        a1, a2 = grid.basis[6:9], grid.basis[9:12]
        nl, nc = grid.nlev, grid.ncol
        rho = np.empty((nl, nc))
        u1 = np.empty((nl, nc))
        u2 = np.empty((nl, nc))
        rhoe = np.empty((nl, nc))
        for k in range(nl):
            z = zc[k]
            s = grid.scale(z)
            r = self.rho(n, z)
            V = self.V(n, z)
            rho[k] = r
            u1[k] = s * np.einsum("in,in->n", V, a1)
            u2[k] = s * np.einsum("in,in->n", V, a2)
            # K of the discrete state uses the centre-interpolated u_3, as the
            # reference's compute_kinetic does; the energy density is set with
            # the analytic K at the centre (the difference is second order in dz)
            e_int = P.cv_d * (self.T(n, z) - P.T_0) - P.R_d * P.T_0
            rhoe[k] = r * (e_int + self.K(n, z) + P.grav * z)
        u3 = np.empty((nl + 1, nc))
        for f in range(nl + 1):
            u3[f] = Jvf[f] * self.w(n, zf[f])
        return np.concatenate([a.reshape(-1) for a in (rho, u1, u2, rhoe, u3)])

    # ------------------------------------------- analytic explicit tendency
    def _u3d(self, x, R):
        """Full 3-D velocity V + w n at Cartesian points x (3, N)."""
        rr = np.sqrt((x * x).sum(0))
        n = x / rr[None, :]
        return self.V(n, rr - R) + self.w(n, rr - R)[None, :] * n

    def _K3d(self, x, R):
        rr = np.sqrt((x * x).sum(0))
        return self.K(x / rr[None, :], rr - R)

    def _rotational(self, n, z, R):
        """-(omega + 2 Omega) x u - grad K at height z (3, N): omega = curl u
        and grad K from central differences of step 1e-2 m in x."""
        Om = np.array([0.0, 0.0, self.p.Omega])[:, None]
        x = (R + z) * n
        dx = 1e-2
        Jx = np.empty((3, 3, n.shape[1]))
        gK = np.empty((3, n.shape[1]))
        for kk in range(3):
            e = np.zeros((3, 1))
            e[kk] = dx
            Jx[:, kk] = (self._u3d(x + e, R) - self._u3d(x - e, R)) / (2 * dx)
            gK[kk] = (self._K3d(x + e, R) - self._K3d(x - e, R)) / (2 * dx)
        omega = np.stack([Jx[2, 1] - Jx[1, 2], Jx[0, 2] - Jx[2, 0], Jx[1, 0] - Jx[0, 1]])
        u = self.V(n, z) + self.w(n, z)[None, :] * n
        return -cross(omega + 2.0 * Om, u) - gK

    def explicit_tendency(self, grid):
        """Analytic explicit tendencies: (rho, u1, u2, rhoe) at centres, each
        (nlev, ncol), and u_3 at faces (nlev + 1, ncol; the two boundary faces,
        where the impenetrability filter acts, are left 0)."""
        n = unit_vectors(grid)
        a1, a2 = grid.basis[6:9], grid.basis[9:12]
        R = grid.radius
        _, _, Jvf = grid.vertical()
        out = [np.empty((grid.nlev, grid.ncol)) for _ in range(4)]
        for k, z in enumerate(grid.zc):
            r = R + z
            s = grid.scale(z)
            out[0][k] = -surf_div(lambda m: self.rho(m, z) * self.V(m, z), n, r)
            out[3][k] = -surf_div(lambda m: self.rho(m, z) * self.h_tot(m, z) * self.V(m, z), n, r)
            gp = surf_grad(lambda m: self.p_of(m, z), n, r)
            Vt = proj(n, self._rotational(n, z, R)) - gp / self.rho(n, z)[None, :]
            out[1][k] = s * np.einsum("in,in->n", Vt, a1)
            out[2][k] = s * np.einsum("in,in->n", Vt, a2)
        u3 = np.zeros((grid.nlev + 1, grid.ncol))
        for f, z in enumerate(grid.z_face):
            if 0 < f < grid.nlev:
                u3[f] = Jvf[f] * (self._rotational(n, z, R) * n).sum(0)
        return out + [u3]


# ------------------------------------------------------------ hyperdiffusion
# Degree-3 harmonic homogeneous polynomial: its restriction to the unit sphere
# is a spherical harmonic of degree l = 3, an eigenfunction of the surface
# Laplacian with eigenvalue -l (l + 1) / r^2 on the sphere of radius r.
L_DEG = 3


def harmonic(n):
    return n[0] * n[1] * n[2] + 0.5 * (n[0] ** 3 - 3.0 * n[0] * n[1] ** 2)


def grad_harmonic(n):
    return np.stack([n[1] * n[2] + 1.5 * (n[0] ** 2 - n[1] ** 2), n[0] * n[2] - 3.0 * n[0] * n[1],
                     n[0] * n[1]])


def harmonic_state(grid, params, kind, amp=10.0):
    """States whose hyperdiffusion tendency is known in closed form.

    kind = "rot": u_h = amp (n x grad Y)      (rotational; curl-curl part)
           "div": u_h = amp P grad Y          (divergent; damped by delta)
           "scalar": p varies horizontally and T is set so that
                     s_d - s_d,r(p) = const + cp_d amp Y (rho varies)
           "w": u_3 = Jv amp Y (uniform vertical grid)
    Returns (Y, info) with the fields the closed form needs."""
    P = params
    n = unit_vectors(grid)
    a1, a2 = grid.basis[6:9], grid.basis[9:12]
    nl, nc = grid.nlev, grid.ncol
    zc = grid.zc
    _, _, Jvf = grid.vertical()
    Yh = harmonic(n)
    rho = np.empty((nl, nc))
    u1 = np.zeros((nl, nc))
    u2 = np.zeros((nl, nc))
    rhoe = np.empty((nl, nc))
    u3 = np.zeros((nl + 1, nc))
    for k, z in enumerate(zc):
        s = grid.scale(z)
        p0 = np.full(nc, 1.0e5 * np.exp(-z / 8000.0))
        T = np.full(nc, 260.0 - 0.002 * z)
        V = np.zeros((3, nc))
        if kind == "rot":
            V = amp * cross(n, grad_harmonic(n))
        elif kind == "div":
            V = amp * proj(n, grad_harmonic(n))
        elif kind == "scalar":
            p0, T = scalar_harmonic_pT(n, z, params, amp)
        u1[k] = s * np.einsum("in,in->n", V, a1)
        u2[k] = s * np.einsum("in,in->n", V, a2)
        r = p0 / (P.R_d * T)
        rho[k] = r
        K = 0.5 * np.einsum("in,in->n", V, V)
        rhoe[k] = r * (P.cv_d * (T - P.T_0) - P.R_d * P.T_0 + K + P.grav * z)
    if kind == "w":
        for f in range(1, nl):
            u3[f] = Jvf[f] * amp * Yh
    Y = np.concatenate([a.reshape(-1) for a in (rho, u1, u2, rhoe, u3)])
    return Y, dict(rho=rho, u1=u1, u2=u2, u3=u3, n=n)


def _refstate(params, p):
    """T_r(p) and Phi_r(p) of refstate_thermodynamics.jl:80-88 and :123-131
    (s_ref = 7, Pi = (p / p_ref_theta)^(R_d / cp_d))."""
    Pi = (p / params.p_ref_theta) ** (params.R_d / params.cp_d)
    Tr = params.T_min_ref + (params.T_surf_ref - params.T_min_ref) * Pi ** 7
    Phr = -params.cp_d * (params.T_min_ref * np.log(Pi) + (params.T_surf_ref - params.T_min_ref) / 7 * (Pi ** 7 - 1))
    return Tr, Phr


def scalar_harmonic_pT(n, z, params, amp):
    """Pressure varying across the level and the temperature for which the
    hyperdiffused scalar s_d - s_d,r(p) = cp_d (T - T_r(p)) + g z - Phi_r(p)
    equals cp_d amp Y(n) + const (sd_r, refstate_thermodynamics.jl:161-166)."""
    p = 1.0e5 * np.exp(-z / 8000.0) * (1.0 + 0.02 * (n[0] - 0.5 * n[1] * n[2]))
    Tr, Phr = _refstate(params, p)
    T = Tr + (Phr - params.grav * z) / params.cp_d + amp * harmonic(n)
    return p, T


def hyperdiffusion_closed_form(grid, params, kind, info, nu_v, nu_s, delta, amp=10.0):
    """Closed-form hyperdiffusion tendency of harmonic_state(kind): the field
    it acts on (name, plane array) and its expected tendency."""
    lam = lambda r: L_DEG * (L_DEG + 1) / r ** 2
    R = grid.radius
    if kind in ("rot", "div"):
        fac = np.array([lam(R + z) ** 2 for z in grid.zc])[:, None]
        d = delta if kind == "div" else 1.0
        return (-nu_v * d * fac * info["u1"], -nu_v * d * fac * info["u2"])
    if kind == "w":
        fac = np.array([lam(R + z) ** 2 for z in grid.z_face])[:, None]
        return -nu_v * fac * info["u3"]
    if kind == "scalar":
        # rho e_t = -nu_s wdiv(rho grad L), L = lap(s_d) = -lam cp_d amp Y
        n = info["n"]
        out = np.empty((grid.nlev, grid.ncol))
        for k, z in enumerate(grid.zc):
            r = R + z

            def flux(m):
                mm = m / np.sqrt((m * m).sum(0))[None, :]
                p, T = scalar_harmonic_pT(m, z, params, amp)
                rho = p / (params.R_d * T)
                gL = -lam(r) * params.cp_d * amp * proj(mm, grad_harmonic(m)) / r
                return rho[None, :] * gL

            out[k] = -nu_s * surf_div(flux, n, r)
        return out
    raise ValueError(kind)


# -------------------------------------------------------- vertical diffusion
def vdiff_tendency(sm, grid, D0):
    """Analytic vertical-diffusion tendencies of the Smooth state sm (w = 0) on
    the deep sphere with a constant diffusivity D0 (vertical_diffusion_boundary_
    layer.jl:63-103 with DecayWithHeightDiffusion(H = Inf)):

        rhoe_t = r^-2 d/dr (r^2 rho D0 d s_d / dr),   s_d = cp_d (T - T_0) + g z
        V_t    = (rho r^2)^-1 d/dr (r^2 rho D0 dV / dr)

    at the centres (interior levels are returned; the two boundary levels see
    the zero-flux boundary condition).  Derivatives are nested central
    differences of step 0.5 m."""
    P = sm.p
    n = unit_vectors(grid)
    a1, a2 = grid.basis[6:9], grid.basis[9:12]
    R = grid.radius
    dz = 0.5

    def s_d(z):
        return P.cp_d * (sm.T(n, z) - P.T_0) + P.grav * z

    def d(f, z):
        return (f(z + dz) - f(z - dz)) / (2 * dz)

    def flux_s(z):
        return (R + z) ** 2 * sm.rho(n, z) * D0 * d(s_d, z)

    def flux_V(z):
        return (R + z) ** 2 * sm.rho(n, z)[None, :] * D0 * d(lambda q: sm.V(n, q), z)

    out_e = np.zeros((grid.nlev, grid.ncol))
    out_1 = np.zeros((grid.nlev, grid.ncol))
    out_2 = np.zeros((grid.nlev, grid.ncol))
    for k, z in enumerate(grid.zc):
        if k == 0 or k == grid.nlev - 1:
            continue
        r = R + z
        out_e[k] = d(flux_s, z) / r ** 2
        Vt = d(flux_V, z) / (r ** 2 * sm.rho(n, z)[None, :])
        s = grid.scale(z)
        out_1[k] = s * np.einsum("in,in->n", Vt, a1)
        out_2[k] = s * np.einsum("in,in->n", Vt, a2)
    return out_e, out_1, out_2
