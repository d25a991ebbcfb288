"""Seeded initial states of the dry-dycore workloads.

The state is the reference's prognostic set (``src/setups/common/
prognostic_variables.jl:21-61,280-287``): ``Y.c = (rho, u_h (covariant C12),
rho e_tot)`` at centres and ``Y.f = u_3`` (covariant C3) at faces, with
``rho e_tot = rho (e_int(T) + K + Phi)``.  The dry internal energy is
``e_int = cv_d (T - T_0) - R_d T_0`` (the form implied by the Jacobian's
``dp/drho`` in ``implicit/manual_sparse_jacobian.jl:816-818``) and
``Phi = g z`` (``utils/utilities.jl:84``).

Native flat layout (float, level-major): ``[rho | u1 | u2 | rho_e | u3]``,
each centre plane ``(nlev, ncol)`` and the face plane ``(nlev + 1, ncol)``.

Noise is drawn once with numpy's PCG64 generator per *unique* horizontal node,
so columns duplicated at element edges start bitwise identical.
"""
from __future__ import annotations

import numpy as np

from .geometry import Grid


def balanced_profile(z, params, T_s=300.0, lapse=6.5e-3, z_trop=12000.0, p_s=1.0e5):
    """Dry hydrostatic profile: ``T = T_s - lapse z`` up to ``z_trop``, then
    isothermal; ``p`` from the closed-form integral of ``dp/dz = -rho g``."""
    z = np.asarray(z, dtype=np.float64)
    g, R = params.grav, params.R_d
    T_trop = T_s - lapse * z_trop
    T = np.where(z <= z_trop, T_s - lapse * z, T_trop)
    expo = g / (R * lapse)
    p_trop = p_s * (T_trop / T_s) ** expo
    p_lo = p_s * (np.maximum(T_s - lapse * z, 1.0) / T_s) ** expo
    p_hi = p_trop * np.exp(-g * (z - z_trop) / (R * T_trop))
    p = np.where(z <= z_trop, p_lo, p_hi)
    return T, p


def e_int_dry(T, params):
    return params.cv_d * (T - params.T_0) - params.R_d * params.T_0


def layout_sizes(ncol: int, nlev: int):
    nc = nlev * ncol
    return nc, (nlev + 1) * ncol, 4 * nc + (nlev + 1) * ncol


def pack(rho, u1, u2, rhoe, u3):
    return np.concatenate([a.reshape(-1) for a in (rho, u1, u2, rhoe, u3)])


def unpack(Y, ncol: int, nlev: int):
    nc = nlev * ncol
    Y = np.asarray(Y)
    return (Y[0:nc].reshape(nlev, ncol), Y[nc:2 * nc].reshape(nlev, ncol),
            Y[2 * nc:3 * nc].reshape(nlev, ncol), Y[3 * nc:4 * nc].reshape(nlev, ncol),
            Y[4 * nc:].reshape(nlev + 1, ncol))


FIELD_NAMES = ("rho", "u1", "u2", "rhoe", "u3")


def _assemble(grid: Grid, params, T, p, u_e, u_n):
    """State from temperature/pressure (nlev, ncol) and physical wind."""
    zc = grid.zc[:, None]
    rho = p / (params.R_d * T)
    s = grid.scale(zc)
    b = grid.basis
    if grid.kind == "sphere":
        lat, lon = grid.lat, grid.lon
        ee = np.stack([-np.sin(lon), np.cos(lon), np.zeros_like(lon)])
        en = np.stack([-np.sin(lat) * np.cos(lon), -np.sin(lat) * np.sin(lon), np.cos(lat)])
        a1, a2 = b[6:9], b[9:12]
        # covariant components u_i = u_phys . a_i(z), a_i(z) = s a_i(R)
        ue1 = np.einsum("ic,ic->c", ee, a1)
        ue2 = np.einsum("ic,ic->c", ee, a2)
        un1 = np.einsum("ic,ic->c", en, a1)
        un2 = np.einsum("ic,ic->c", en, a2)
        u1 = s * (u_e * ue1 + u_n * un1)
        u2 = s * (u_e * ue2 + u_n * un2)
    else:
        u1 = u_e * np.ones_like(T)
        u2 = u_n * np.ones_like(T)
    K = 0.5 * (u_e ** 2 + u_n ** 2)
    Phi = params.grav * zc
    rhoe = rho * (e_int_dry(T, params) + K + Phi)
    u3 = np.zeros((grid.nlev + 1, grid.ncol))
    return rho, u1, u2, rhoe, u3


def column_state(grid: Grid, params, seed: int = 0, rel_noise: float = 1e-3,
                 u=(10.0, 0.0), T_s=None):
    """Balanced single-column state with ``u_h = (10, 0)``, ``u_3 = 0`` and
    ``rho e_tot`` multiplied by ``1 + rel_noise N(0, 1)``.  ``T_s`` may be an
    array of per-column surface temperatures."""
    zc = grid.zc
    ncol = grid.ncol
    if T_s is None:
        T_s = np.full(ncol, 300.0)
    T = np.empty((grid.nlev, ncol))
    p = np.empty((grid.nlev, ncol))
    for c in range(ncol):
        T[:, c], p[:, c] = balanced_profile(zc, params, T_s=float(T_s[c]))
    rho, u1, u2, rhoe, u3 = _assemble(grid, params, T, p, u[0], u[1])
    rng = np.random.default_rng(seed)
    rhoe = rhoe * (1.0 + rel_noise * rng.standard_normal(rhoe.shape))
    return pack(rho, u1, u2, rhoe, u3)


def sphere_state(grid: Grid, params, seed: int = 1, jet: float = 35.0,
                 noise_K: float = 1.0, noise_top: float = 5000.0):
    """Balanced profile + zonal jet ``u = jet sin^2(2 lat)`` + seeded
    ``noise_K`` Gaussian temperature noise below ``noise_top``; ``p`` is the
    balanced pressure and ``rho = p / (R_d T)``."""
    zc = grid.zc
    T1, p1 = balanced_profile(zc, params)
    T = np.repeat(T1[:, None], grid.ncol, axis=1)
    p = np.repeat(p1[:, None], grid.ncol, axis=1)
    rng = np.random.default_rng(seed)
    nz = int(np.sum(zc < noise_top))
    noise = rng.standard_normal((nz, grid.n_nodes)) * noise_K
    T[:nz, :] += noise[:, grid.node_id]
    u_e = jet * np.sin(2.0 * grid.lat) ** 2
    return pack(*_assemble(grid, params, T, p, u_e[None, :], 0.0))


def surface_temperature(grid: Grid) -> np.ndarray:
    """Per-column surface temperature of the default ``AtmosSurface``:
    ``AnalyticTemperature(zonally_symmetric_temperature)`` (reference
    ``src/setups/Setups.jl:205-216``, ``src/types.jl:2292-2302``), evaluated at
    the surface (z = 0, no topography).  On a sphere ``271 + 29
    exp(-lat^2 / (2 26^2)) - 6.5e-3 z`` with the latitude in degrees; on
    independent columns (no latitude-longitude coordinates) 300 K."""
    if grid.nelem == 0:
        return np.full(grid.ncol, 300.0)
    lat = np.degrees(np.asarray(grid.lat, np.float64))
    return 271.0 + 29.0 * np.exp(-lat ** 2 / (2.0 * 26.0 ** 2)) - 6.5e-3 * 0.0
