"""Physical constants and numerics settings of the dry-dycore step.

Constants come from ``ca_params.toml`` (ClimaParams key names).  Derived values
follow Thermodynamics' dry-air relations: R_d = gas_constant/molar_mass_dry_air,
cp_d = R_d/kappa_d, cv_d = cp_d - R_d (reference ``src/parameters/
create_parameters.jl:202-224`` maps the ClimaParams names onto ClimaAtmos).

Numerics defaults follow ``config/default_configs/default_config.yml``:
hyperdiffusion c_nu = 0.1857 (:30-32), Pr = 0.2 (:33-35), divergence damping 5
(:218-220); energy upwinding ``vanleer_limiter`` (``src/types.jl:1916``).
"""
from __future__ import annotations

import math
import os
from dataclasses import dataclass, field, asdict

try:
    import tomli as _toml
except ImportError:  # pragma: no cover - tomli is in the image
    _toml = None

_HERE = os.path.dirname(os.path.abspath(__file__))
PARAMS_TOML = os.path.join(_HERE, "ca_params.toml")

UPWIND_NONE, UPWIND_FIRST, UPWIND_THIRD, UPWIND_VANLEER = 0, 1, 2, 3
_UPWIND_NAMES = {
    "none": UPWIND_NONE,
    "first_order": UPWIND_FIRST,
    "third_order": UPWIND_THIRD,
    "vanleer_limiter": UPWIND_VANLEER,
}


def _load_toml(path):
    with open(path, "rb") as f:
        raw = _toml.load(f)
    return {k: float(v["value"]) for k, v in raw.items()}


@dataclass
class Params:
    """Thermodynamic and planetary constants (SI units)."""

    R_d: float
    cp_d: float
    cv_d: float
    T_0: float
    p_ref_theta: float
    grav: float
    Omega: float
    T_surf_ref: float
    T_min_ref: float
    T_min_sgs: float
    f_plane: float
    planet_radius: float
    zd_rayleigh: float
    alpha_rayleigh_uh: float
    alpha_rayleigh_w: float

    @classmethod
    def from_toml(cls, path: str = PARAMS_TOML) -> "Params":
        t = _load_toml(path)
        R_d = t["gas_constant"] / t["molar_mass_dry_air"]
        cp_d = R_d / t["adiabatic_exponent_dry_air"]
        return cls(
            R_d=R_d,
            cp_d=cp_d,
            cv_d=cp_d - R_d,
            T_0=t["thermodynamics_temperature_reference"],
            p_ref_theta=t["potential_temp_reference_pressure"],
            grav=t["gravitational_acceleration"],
            Omega=t["angular_velocity_planet_rotation"],
            T_surf_ref=t["temperature_surface_reference"],
            T_min_ref=t["temperature_min_reference"],
            T_min_sgs=t["temperature_minimum"],
            f_plane=t["f_plane_coriolis_frequency"],
            planet_radius=t["planet_radius"],
            zd_rayleigh=t["zd_rayleigh"],
            alpha_rayleigh_uh=t["alpha_rayleigh_uh"],
            alpha_rayleigh_w=t["alpha_rayleigh_w"],
        )


@dataclass
class Numerics:
    """Numerics of the parity profile.

    ``nu4_vorticity = c_nu * h**3`` and ``nu4_scalar = nu4_vorticity / Pr``
    (reference ``src/prognostic_equations/hyperdiffusion.jl:21-28``), with
    ``h`` the mean nodal spacing of the horizontal grid.
    """

    hyperdiff: bool = True
    c_nu: float = 0.1857
    prandtl: float = 0.2
    div_damping: float = 5.0
    upwinding: int = UPWIND_VANLEER
    rayleigh: bool = False
    vert_diff: bool = False
    deep: bool = True
    dt: float = 100.0
    # surface_flux_tendency! with the default AtmosSurface (src/types.jl:2292-2302):
    # ExchangeCoefficients(Cd = Ch = 0.0044), gustiness 1 m/s (surface_conditions.jl:254)
    # vertical diffusion in the implicit tendency (diff_mode Implicit, YAML
    # implicit_diffusion) and the iterations of the approximate linear solve it
    # brings in (approximate_linear_solve_iters, default 1)
    implicit_diffusion: bool = False
    solve_iters: int = 1
    sfc_flux: bool = False
    Cd: float = 0.0044
    Ch: float = 0.0044
    gustiness: float = 1.0
    # the vertical-diffusion closure (vert_diff, src/config/model_getters.jl:341-355):
    # "DecayWithHeightDiffusion" (K = D0_col exp(-z / H); H = inf keeps K = D0_col)
    # or "VerticalDiffusion" (K = C_E |u_h(lowest centre)| dz_1 / 2, tapered above
    # 850 hPa; precomputed_quantities.jl:651-676)
    vert_diff_model: str = "DecayWithHeightDiffusion"
    vert_diff_H: float = math.inf
    vert_diff_C_E: float = 0.044

    @staticmethod
    def upwinding_code(name: str) -> int:
        return _UPWIND_NAMES[name]

    def vert_diff_kind(self) -> int:
        return {"DecayWithHeightDiffusion": 0, "VerticalDiffusion": 1}[self.vert_diff_model]

    def nu4(self, h: float):
        nu_v = self.c_nu * h**3
        return nu_v, nu_v / self.prandtl


def node_horizontal_length_scale(radius: float, h_elem: int, nq: int = 4) -> float:
    """Mean nodal spacing sqrt(area / (n_elem (nq-1)^2)) of a cubed sphere.

    For the sphere area 4 pi R^2 and 6 h_elem^2 elements this is
    sqrt(4 pi / 6) R / ((nq-1) h_elem), the length used by the CAM-SE rescale
    factor in reference ``src/types.jl:506-516``.
    """
    return math.sqrt(4.0 * math.pi / 6.0) * radius / ((nq - 1) * h_elem)
