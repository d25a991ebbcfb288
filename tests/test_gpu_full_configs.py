"""Size-independent properties of the step with the optional terms on:
surface fluxes (surface_flux_tendency!, surface_flux.jl:101-145) and vertical
diffusion treated implicitly (implicit_tendency.jl:73-82 with the approximate
arrowhead solve, manual_sparse_jacobian.jl:538-578).  Neither moves mass, so
the global mass integral is conserved to float32 round-off; the state stays
finite, u_3 stays zero on the boundary faces, and the end-of-step dss! leaves
the copies of a shared node identical."""
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


def _mass(g, zf, Y):
    _, w, _ = G.gll()
    zc, Jvc, _ = G.vertical_metrics(zf)
    s = g.scale(zc)
    J = (np.tile(np.outer(w, w).reshape(-1), g.nelem) * g.J_h)[None, :] * (s ** 2 * Jvc)[:, None]
    rho = I.unpack(Y.astype(np.float64), g.ncol, g.nlev)[0]
    return float((rho * J).sum())


def _run(h_elem, z_elem, dz_bottom, dt, nsteps, seed):
    torch = _torch()
    from cadycore.lib import DryDycore

    zf = G.tanh_stretched_faces(z_elem, 30000.0, dz_bottom)
    g = G.cubed_sphere(h_elem, zf, PAR.planet_radius, PAR.Omega)
    num = P.Numerics(dt=dt, sfc_flux=True, vert_diff=True, implicit_diffusion=True)
    rng = np.random.default_rng(seed)
    D0 = np.exp(rng.uniform(np.log(0.5), np.log(20.0), g.n_nodes))[g.node_id]
    d = DryDycore(g, PAR, num, D0_col=D0, T_sfc_col=I.surface_temperature(g))
    Y = d.state(I.sphere_state(g, PAR, seed=seed))
    d.dss(Y)
    torch.cuda.synchronize()
    m0 = _mass(g, zf, Y.cpu().numpy())
    rhoe0 = I.unpack(Y.cpu().numpy().astype(np.float64), g.ncol, g.nlev)[3].copy()
    d.step(Y, dt, nsteps)
    torch.cuda.synchronize()
    assert d.check_finite(Y)
    Yn = Y.cpu().numpy()
    d.close()
    m1 = _mass(g, zf, Yn)
    rho, u1, u2, rhoe, u3 = I.unpack(Yn, g.ncol, g.nlev)
    assert np.all(u3[0] == 0) and np.all(u3[-1] == 0)
    ptr, cols = g.dss_ptr, g.dss_col
    for n in range(0, len(ptr) - 1, 53):
        members = cols[ptr[n]:ptr[n + 1]]
        for arr in (rho, rhoe, u3):
            col = arr[:, members]
            assert np.all(col == col[:, :1])
    # the surface fluxes do change the lowest layer's energy
    assert np.abs(rhoe[0] - rhoe0[0]).max() > 0
    return m0, m1


def test_full_size_with_surface_fluxes_and_implicit_diffusion_h30():
    m0, m1 = _run(30, 63, 30.0, 60.0, 5, 2)
    print("h30 z63, 5 steps: relative mass change", m1 / m0 - 1)
    assert abs(m1 / m0 - 1) < 2e-6


def test_long_run_with_surface_fluxes_and_implicit_diffusion_h4():
    """300 steps of 100 s (8.3 hours) at h_elem 4, z_elem 16."""
    m0, m1 = _run(4, 16, 100.0, 100.0, 300, 5)
    print("h4 z16, 300 steps: relative mass change", m1 / m0 - 1)
    assert abs(m1 / m0 - 1) < 2e-5
