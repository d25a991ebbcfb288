"""Read-before-write checks of the native library through its debug poisoning
(the counterpart of fill_with_nans!, src/utils/debug_utils.jl:67-70): with
poisoning on, every entry point fills the work buffers no call carries over
with NaN first, so any value a call reads without writing it comes out NaN.
Each result must be bitwise identical to the unpoisoned one.
"""
import numpy as np
import pytest

from cadycore import geometry as G
from cadycore import initial as I
from cadycore import params as P
from cadycore.lib import CA_TUNE_NEWTON

pytestmark = pytest.mark.gpu
PAR = P.Params.from_toml()


def _torch():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _same(a, b):
    a, b = a.cpu().numpy(), b.cpu().numpy()
    return np.array_equal(a, b) and np.all(np.isfinite(a))


def _sphere(nlev=10, vert_diff=True):
    g = G.cubed_sphere(2, G.tanh_stretched_faces(nlev, 30000.0, 300.0), PAR.planet_radius, PAR.Omega)
    num = P.Numerics(dt=100.0, vert_diff=vert_diff)
    D0 = np.exp(np.random.default_rng(3).uniform(np.log(0.1), np.log(50.0), g.n_nodes))[g.node_id]
    Y0 = I.sphere_state(g, PAR, seed=2).astype(np.float32).astype(np.float64)
    return g, num, D0, Y0


def _hooks(d, Y0):
    """Every hook once on the same input; returns the outputs."""
    out = []
    Y = d.state(Y0)
    d.set_implicit_precomputed_quantities(Y)
    out.append(Y.clone())
    Yt = d.zeros()
    d.remaining_tendency(Yt, Y)
    out.append(Yt.clone())
    Yi = d.zeros()
    d.implicit_tendency(Yi, Y)
    out.append(Yi)
    d.update_jacobian(Y, 7.0)
    dY = d.zeros()
    d.ldiv(dY, Yt)
    out.append(dY)
    Yc = d.zeros()
    d.correct_implicit_advection_tendency(Yc, Y)
    out.append(Yc)
    Yd = d.state(Y0)
    d.dss(Yd)
    out.append(Yd)
    Ys = d.state(Y0)
    d.step(Ys, 100.0, 3)
    out.append(Ys)
    return out


@pytest.mark.parametrize("newton", [0, 1])
def test_hooks_and_step_unchanged_under_poisoning(newton):
    torch = _torch()
    from cadycore.lib import DryDycore

    g, num, D0, Y0 = _sphere()
    d = DryDycore(g, PAR, num, D0_col=D0)
    d.set_tuning(CA_TUNE_NEWTON, newton)
    clean = _hooks(d, Y0)
    d.set_debug_poison(True)
    dirty = _hooks(d, Y0)
    torch.cuda.synchronize()
    for i, (a, b) in enumerate(zip(clean, dirty)):
        assert _same(a, b), f"output {i} differs under poisoning"
    d.close()


def test_columns_unchanged_under_poisoning():
    torch = _torch()
    from cadycore.lib import DryDycore

    g = G.columns(5, G.uniform_faces(30, 30000.0), f_plane=1e-4)
    num = P.Numerics(dt=10.0, hyperdiff=False, vert_diff=True, rayleigh=True)
    Y0 = I.column_state(g, PAR, seed=0).astype(np.float32).astype(np.float64)
    d = DryDycore(g, PAR, num, D0_col=np.full(5, 3.0))
    clean = _hooks(d, Y0)
    d.set_debug_poison(True)
    dirty = _hooks(d, Y0)
    torch.cuda.synchronize()
    for i, (a, b) in enumerate(zip(clean, dirty)):
        assert _same(a, b), f"output {i} differs under poisoning"
    d.close()


@pytest.mark.parametrize("nlev,newton", [(10, 1), (10, 0), (40, 1)])
def test_pullback_unchanged_under_poisoning(nlev, newton):
    """The sphere pullback with checkpoints every 2 and every 3 steps, each
    with and without poisoning: all four bitwise identical."""
    torch = _torch()
    from cadycore.lib import DryDycore

    g, num, D0, Y0 = _sphere(nlev)
    rng = np.random.default_rng(5)
    w = (rng.standard_normal(len(Y0)) / np.maximum(np.abs(Y0), 1.0)).astype(np.float32)
    d = DryDycore(g, PAR, num, D0_col=D0)
    d.set_tuning(CA_TUNE_NEWTON, newton)
    res = []
    for poison in (False, True):
        d.set_debug_poison(poison)
        for every in (2, 3):
            D0b = torch.zeros(g.ncol, dtype=torch.float32, device="cuda")
            Yb = d.step_adjoint(d.state(Y0), d.state(w), num.dt, 4, every=every, D0b=D0b)
            torch.cuda.synchronize()
            res.append((Yb.clone(), D0b.clone()))
    for i in range(1, 4):
        assert _same(res[0][0], res[i][0]) and _same(res[0][1], res[i][1]), f"pullback {i} differs"
    d.close()


def test_reverse_mode_rejects_more_than_128_levels():
    """The sphere reverse mode keeps a column's faces in one workgroup's LDS
    rows of 129 entries: deeper grids get CA_ERR_ARG before any launch."""
    _torch()
    from cadycore.lib import DryDycore

    g = G.cubed_sphere(2, G.uniform_faces(130, 30000.0), PAR.planet_radius, PAR.Omega)
    d = DryDycore(g, PAR, P.Numerics(dt=10.0))
    Y = d.zeros()
    with pytest.raises(RuntimeError, match="nlev <= 128"):
        d.step_adjoint(Y, d.zeros(), 10.0, 2, every=1)
    with pytest.raises(RuntimeError, match="nlev <= 128"):
        d.step_record(Y, 10.0, 2, every=1)
    d.close()
