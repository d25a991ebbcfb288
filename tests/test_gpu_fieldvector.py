"""The hooks on ClimaCore FieldVectors (ca_*_fv) against the native-pointer
hooks, through the C ABI: every hook called on ClimaCore's VIJFH parent arrays
(Julia (Nv, Ni, Nj, F, Nh), level fastest) must give bitwise the result of
the same hook on the native layout, and the native descriptor must be used in
place.  Reference hooks: remaining_tendency!(Yt, Yt_lim, Y, p, t)
(remaining_tendency.jl:48), implicit_tendency!(Yt, Y, p, t)
(implicit_tendency.jl:36), update_jacobian! and ldiv! (jacobian.jl:74-82),
T_post_imp! (implicit_tendency.jl:322), cache! / cache_imp!
(precomputed_quantities.jl:698, 928) and dss! (constrain_state.jl:59).
"""
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


@pytest.fixture(scope="module")
def dyc():
    _torch()
    from cadycore.lib import DryDycore

    g = G.cubed_sphere(2, G.tanh_stretched_faces(10, 30000.0, 300.0), PAR.planet_radius, PAR.Omega)
    num = P.Numerics(dt=100.0, vert_diff=True)
    rng = np.random.default_rng(3)
    d = DryDycore(g, PAR, num, D0_col=rng.uniform(0.1, 50.0, g.ncol))
    Y0 = I.sphere_state(g, PAR, seed=1)
    # a vertical velocity, and nonzero u_3 on the boundary faces for the filter
    nC = g.nlev * g.ncol
    Y0[4 * nC:] = rng.normal(0.0, 5.0, (g.nlev + 1) * g.ncol)
    d.Y0 = d.state(Y0.astype(np.float32).astype(np.float64))
    yield d
    d.close()


def _same(a, b, what):
    a, b = a.cpu().numpy(), b.cpu().numpy()
    assert np.all(np.isfinite(a)), what
    assert np.array_equal(a, b), f"{what}: {np.count_nonzero(a != b)} differ, max {np.abs(a - b).max():.3e}"


def test_vijfh_round_trip(dyc):
    c, f = dyc.to_vijfh(dyc.Y0)
    _same(dyc.from_vijfh(c, f), dyc.Y0, "VIJFH round trip")


def test_cache_hooks_filter_in_place(dyc):
    torch = _torch()
    Yn = dyc.Y0.clone()
    dyc.set_implicit_precomputed_quantities(Yn)
    for hook in (dyc.set_implicit_precomputed_quantities_fv, dyc.set_precomputed_quantities_fv):
        c, f = dyc.to_vijfh(dyc.Y0)
        hook(dyc.fieldvector(c=c, f=f), 0.0)
        torch.cuda.synchronize()
        _same(dyc.from_vijfh(c, f), Yn, hook.__name__)
    Y2 = dyc.Y0.clone()
    dyc.set_precomputed_quantities_fv(dyc.fieldvector(Y=Y2))
    torch.cuda.synchronize()
    _same(Y2, Yn, "native descriptor")


def test_tendency_hooks_on_vijfh(dyc):
    torch = _torch()
    Y = dyc.Y0.clone()
    dyc.set_implicit_precomputed_quantities(Y)
    c, f = dyc.to_vijfh(Y)
    Yfv = dyc.fieldvector(c=c, f=f)
    for name in ("remaining_tendency", "implicit_tendency", "correct_implicit_advection_tendency"):
        ref = dyc.zeros()
        getattr(dyc, name)(ref, Y)
        tc, tf = (torch.full_like(x, 3.0) for x in (c, f))
        Ytfv = dyc.fieldvector(c=tc, f=tf)
        if name == "remaining_tendency":
            lc, lf = (torch.full_like(x, 7.0) for x in (c, f))
            dyc.remaining_tendency_fv(Ytfv, dyc.fieldvector(c=lc, f=lf), Yfv, 0.0)
        else:
            getattr(dyc, name + "_fv")(Ytfv, Yfv, 0.0)
        torch.cuda.synchronize()
        _same(dyc.from_vijfh(tc, tf), ref, name)
        if name == "remaining_tendency":
            assert torch.all(lc == 0) and torch.all(lf == 0), "Yt_lim must be zeroed"
        # mixed: native Yt, VIJFH Y
        out = torch.full_like(ref, 5.0)
        if name == "remaining_tendency":
            dyc.remaining_tendency_fv(dyc.fieldvector(Y=out), None, Yfv, 0.0)
        else:
            getattr(dyc, name + "_fv")(dyc.fieldvector(Y=out), Yfv, 0.0)
        torch.cuda.synchronize()
        _same(out, ref, name + " (native Yt)")


def test_jacobian_and_ldiv_on_vijfh(dyc):
    torch = _torch()
    Y = dyc.Y0.clone()
    dyc.set_implicit_precomputed_quantities(Y)
    T = dyc.zeros()
    dyc.implicit_tendency(T, Y)
    dtg = 0.4358665215 * 100.0
    R = (dtg * T).contiguous()
    dyc.update_jacobian(Y, dtg)
    ref = dyc.zeros()
    dyc.ldiv(ref, R)
    torch.cuda.synchronize()
    c, f = dyc.to_vijfh(Y)
    dyc.update_jacobian_fv(dyc.fieldvector(c=c, f=f), dtg, 0.0)
    rc, rf = dyc.to_vijfh(R)
    oc, of = (torch.full_like(x, 9.0) for x in (rc, rf))
    dyc.ldiv_fv(dyc.fieldvector(c=oc, f=of), dyc.fieldvector(c=rc, f=rf))
    torch.cuda.synchronize()
    _same(dyc.from_vijfh(oc, of), ref, "ldiv")


def test_dss_on_vijfh(dyc):
    torch = _torch()
    Yn = dyc.Y0.clone()
    dyc.dss(Yn)
    c, f = dyc.to_vijfh(dyc.Y0)
    dyc.dss_fv(dyc.fieldvector(c=c, f=f), 0.0)
    torch.cuda.synchronize()
    _same(dyc.from_vijfh(c, f), Yn, "dss")
