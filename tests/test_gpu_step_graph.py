"""ca_step_graph: the fused step captured once in a hipGraph and replayed.
Every replay must give bitwise the state of the eager ca_step, also with the
program's own eager work and synchronisations between the replays, after a
change of dt, state buffer or launch schedule (which capture again), and for
the configurations whose stage is another kernel (implicit diffusion with
surface fluxes, independent columns)."""
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


def _eager(d, Y0, dt, n):
    Y = d.state(Y0)
    d.step(Y, dt, n)
    return Y


def _check_replays(d, Y0, dt):
    """Replays against eager ca_step calls of the same lengths: every call
    starts with dss! of its state, which is exact only up to rounding, so
    the call boundaries are part of the result."""
    torch = _torch()
    side = torch.cuda.Stream()
    Yg, Ye = d.state(Y0), d.state(Y0)
    torch.cuda.synchronize()
    done = 0
    for n in (1, 1, 2, 1):
        d.step_graph(Yg, dt, n, side)
        side.synchronize()
        done += n
        # eager work of the program between the replays
        x = torch.randn(1 << 16, device=Yg.device)
        float((x * x).sum())
        d.step(Ye, dt, n)
        torch.cuda.synchronize()
        assert torch.equal(Yg, Ye), f"replay after {done} steps differs"
    return side, Yg, Ye


def test_sphere_replays_bitwise():
    torch = _torch()
    from cadycore.lib import CA_TUNE_NEWTON, DryDycore

    g = G.cubed_sphere(4, G.tanh_stretched_faces(16, 30000.0, 100.0), PAR.planet_radius, PAR.Omega)
    d = DryDycore(g, PAR, P.Numerics(dt=100.0))
    try:
        Y0 = I.sphere_state(g, PAR, seed=6)
        side, Yg, Ye = _check_replays(d, Y0, 100.0)
        # another dt: captured again
        d.step_graph(Yg, 50.0, 2, side)
        side.synchronize()
        d.step(Ye, 50.0, 2)
        torch.cuda.synchronize()
        assert torch.equal(Yg, Ye)
        # another state buffer: captured again
        Y2 = d.state(Y0)
        torch.cuda.synchronize()
        d.step_graph(Y2, 100.0, 3, side)
        side.synchronize()
        assert torch.equal(Y2, _eager(d, Y0, 100.0, 3))
        # another Newton schedule: the tuning call drops the captured step
        d.set_tuning(CA_TUNE_NEWTON, 0)
        Y3 = d.state(Y0)
        torch.cuda.synchronize()
        d.step_graph(Y3, 100.0, 2, side)
        side.synchronize()
        assert torch.equal(Y3, _eager(d, Y0, 100.0, 2))
    finally:
        d.close()


def test_implicit_diffusion_and_surface_fluxes_replay_bitwise():
    from cadycore.lib import DryDycore

    g = G.cubed_sphere(4, G.tanh_stretched_faces(16, 30000.0, 100.0), PAR.planet_radius, PAR.Omega)
    rng = np.random.default_rng(2)
    D0 = np.exp(rng.uniform(np.log(0.5), np.log(20.0), g.n_nodes))[g.node_id]
    num = P.Numerics(dt=100.0, sfc_flux=True, vert_diff=True, implicit_diffusion=True)
    d = DryDycore(g, PAR, num, D0_col=D0, T_sfc_col=I.surface_temperature(g))
    try:
        _check_replays(d, I.sphere_state(g, PAR, seed=7), 100.0)
    finally:
        d.close()


def test_columns_replay_bitwise():
    from cadycore.lib import DryDycore

    g = G.columns(64, G.tanh_stretched_faces(30, 30000.0, 500.0))
    d = DryDycore(g, PAR, P.Numerics(hyperdiff=False, dt=10.0))
    try:
        _check_replays(d, I.column_state(g, PAR, seed=1), 10.0)
    finally:
        d.close()


def test_graph_needs_a_stream_and_no_profiling():
    torch = _torch()
    import ctypes as C

    from cadycore.lib import DryDycore

    g = G.cubed_sphere(2, G.tanh_stretched_faces(10, 30000.0, 300.0), PAR.planet_radius, PAR.Omega)
    d = DryDycore(g, PAR, P.Numerics(dt=100.0))
    try:
        Y = d.state(I.sphere_state(g, PAR, seed=1))
        assert d.lib.ca_step_graph(d.h, d._ptr(Y), 100.0, 1, None) == 1
        assert b"non-default stream" in d.lib.ca_last_error(d.h)
        d.profile_enable(True)
        with pytest.raises(RuntimeError, match="profiling"):
            d.step_graph(Y, 100.0, 1, torch.cuda.Stream())
        d.profile_enable(False)
    finally:
        d.close()
