"""Implicit vertical diffusion (diff_mode Implicit: the vertical diffusion in
implicit_tendency!, its Jacobian blocks and the approximate block arrowhead
solve) -- the native hooks and step against the oracle through the C ABI,
on a column and on the h4 z16 deep sphere, with per-column D0.

Gates as for the other parity tests (relative L2 per field, native float32 vs
oracle float64): rho, u_h, rho e_tot <= 1e-5 for the hooks and one step,
<= 1e-3 after 20 column steps; u_3 is reported against its stated gate.  The
treatment must matter: the implicit and explicit steps differ by far more
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


def errors(Yg, Yo, g, interior_u3=False):
    fg, fo = I.unpack(Yg, g.ncol, g.nlev), I.unpack(Yo, g.ncol, g.nlev)
    out = {}
    for n, a, b in zip(I.FIELD_NAMES, fg, fo):
        if n == "u3" and interior_u3:  # the library returns 0 on the boundary faces
            a, b = a[1:-1], b[1:-1]
        out[n] = rel(a, b) if np.linalg.norm(b) > 0 else float(np.abs(a).max())
    return out


def _case(kind, iters=1, implicit=True):
    rng = np.random.default_rng(7)
    if kind == "column":
        g = G.columns(1, G.tanh_stretched_faces(30, 30000.0, 500.0), PAR.f_plane)
        num = P.Numerics(hyperdiff=False, dt=10.0, vert_diff=True, implicit_diffusion=implicit, solve_iters=iters)
        Y = I.column_state(g, PAR, seed=0)
        D0 = np.array([40.0])
    else:
        g = G.cubed_sphere(4, G.tanh_stretched_faces(16, 30000.0, 100.0), PAR.planet_radius, PAR.Omega)
        num = P.Numerics(dt=100.0, vert_diff=True, implicit_diffusion=implicit, solve_iters=iters)
        Y = I.sphere_state(g, PAR, seed=1)
        # vertical shear of u_h (the jet alone is uniform in height, so its
        # diffusion tendency would be a float32 residual)
        nl, nC = g.nlev, g.nlev * g.ncol
        shear = 1.0 + 0.5 * np.sin(np.pi * g.zc / g.z_face[-1])
        for p in (1, 2):
            Y[p * nC:(p + 1) * nC] *= np.repeat(shear, g.ncol)
        D0 = np.exp(rng.uniform(np.log(1.0), np.log(50.0), g.n_nodes))[g.node_id]
    return g, num, Y.astype(np.float32).astype(np.float64), D0


def _pair(g, num, D0):
    _torch()
    from cadycore.lib import DryDycore
    from oracle import Oracle

    return DryDycore(g, PAR, num, D0_col=D0), Oracle(g, PAR, num, D0_col=D0)


@pytest.mark.parametrize("kind,iters", [("column", 1), ("column", 3), ("sphere", 1), ("sphere", 3)])
def test_implicit_diffusion_hooks_match_oracle(kind, iters):
    torch = _torch()
    g, num, Y32, D0 = _case(kind, iters)
    d, o = _pair(g, num, D0)
    Yo = Y32.copy()
    o.set_precomputed(Yo)
    Yg = d.state(Y32)
    d.set_implicit_precomputed_quantities(Yg)
    out = {}
    T = d.zeros()
    d.remaining_tendency(T, Yg)
    torch.cuda.synchronize()
    out["T_exp"] = errors(T.cpu().numpy().astype(np.float64), o.remaining_tendency(Yo), g, interior_u3=True)
    T = d.zeros()
    d.implicit_tendency(T, Yg)
    To = o.implicit_tendency(Yo)
    torch.cuda.synchronize()
    out["T_imp"] = errors(T.cpu().numpy().astype(np.float64), To, g)
    dtg = 0.4358665215 * num.dt
    d.update_jacobian(Yg, dtg)
    o.update_jacobian(Yo, dtg)
    R = (dtg * To).astype(np.float32).astype(np.float64)
    dY = d.zeros()
    d.ldiv(dY, d.state(R))
    torch.cuda.synchronize()
    out["ldiv"] = errors(dY.cpu().numpy().astype(np.float64), o.ldiv(R), g)
    d.close()
    for hook, e in out.items():
        print(kind, iters, hook, {k: f"{v:.2e}" for k, v in e.items()})
        for name in ("rho", "u1", "u2", "rhoe"):
            if kind == "column" and hook == "T_exp" and name in ("u1", "u2", "rhoe"):
                continue  # the column's explicit tendency is Coriolis only on u_h
            assert e[name] <= 1e-5, (hook, name, e)
        assert e["u3"] <= 1e-4, (hook, e)  # regression guard


# newton 1: the fused wavefront-per-column stage (k_newton_idiff_pcr); newton 0:
# the stage as the sequence of hook kernels (thread per column)
@pytest.mark.parametrize("kind,nsteps,newton,iters", [
    ("column", 1, 1, 1), ("column", 20, 1, 1), ("sphere", 1, 1, 1), ("sphere", 5, 1, 1),
    ("column", 20, 0, 1), ("sphere", 5, 0, 1), ("column", 20, 1, 3), ("sphere", 5, 1, 3)])
def test_implicit_diffusion_steps_match_oracle(kind, nsteps, newton, iters):
    torch = _torch()
    from cadycore.lib import CA_TUNE_NEWTON

    g, num, Y32, D0 = _case(kind, iters)
    d, o = _pair(g, num, D0)
    d.set_tuning(CA_TUNE_NEWTON, newton)
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
    print(kind, nsteps, "newton", newton, "iters", iters, {k: f"{v:.2e}" for k, v in e.items()})
    gate = 1e-5 if nsteps == 1 else 1e-3
    for name in ("rho", "u1", "u2", "rhoe"):
        if kind == "column" and name == "u2" and nsteps == 1:
            continue  # u_2 starts at 0 on the f-plane column: a tiny Coriolis increment
        assert e[name] <= gate, (name, e)
    state = "MET" if e["u3"] <= gate else "UNMET"
    print(f"u3 gate [implicit diffusion {kind} {nsteps} steps, newton {newton}, iters {iters}]: "
          f"{e['u3']:.2e} vs stated {gate:.0e}: {state}")
    assert np.all(np.isfinite(got))
    if kind == "column" or newton == 0 or iters != 1:
        return  # at dt = 10 s on 500 m layers the two treatments differ by less than float32 rounding
    # the treatment matters: the explicit-diffusion oracle step differs
    from oracle import Oracle

    g2, num_e, _, _ = _case(kind, implicit=False)
    oe = Oracle(g2, PAR, num_e, D0_col=D0)
    Ye = Y32.copy()
    oe.dss(Ye)
    oe.set_precomputed(Ye)
    oe.step(Ye, num.dt, nsteps)
    nC = g.nlev * g.ncol
    sl = slice(3 * nC, 4 * nC)
    assert np.linalg.norm(Ye[sl] - Yo[sl]) > 10 * np.linalg.norm(got[sl] - Yo[sl])


def test_implicit_diffusion_rejected_by_the_reverse_mode():
    torch = _torch()
    g, num, Y32, D0 = _case("sphere")
    from cadycore.lib import DryDycore

    d = DryDycore(g, PAR, num, D0_col=D0)
    Y = d.state(Y32)
    st = d.lib.ca_step_record(d.h, C.c_void_p(Y.data_ptr()), 100.0, 2, 1, None)
    assert st == 1 and b"implicit diffusion" in d.lib.ca_last_error(d.h)
    torch.cuda.synchronize()
    d.close()


def test_fused_implicit_diffusion_stage_matches_the_hook_kernels():
    """The fused stage (k_newton_idiff_pcr: PCR solves in registers) and the
    hook-kernel stage (Thomas sweeps over scratch planes) are the same
    algebra with the roundings in another order: after 3 steps of the h4 z16
    sphere they agree to 1e-5 relative per field (u_3 to 1e-4)."""
    torch = _torch()
    from cadycore.lib import CA_TUNE_NEWTON, DryDycore

    g, num, Y32, D0 = _case("sphere")
    out = {}
    for v in (0, 1):
        d = DryDycore(g, PAR, num, D0_col=D0)
        d.set_tuning(CA_TUNE_NEWTON, v)
        Y = d.state(Y32)
        d.dss(Y)
        d.step(Y, num.dt, 3)
        torch.cuda.synchronize()
        out[v] = Y.cpu().numpy().astype(np.float64)
        d.close()
    e = errors(out[1], out[0], g)
    print("fused vs hook kernels, 3 steps:", {k: f"{v:.2e}" for k, v in e.items()})
    for name in ("rho", "u1", "u2", "rhoe"):
        assert e[name] <= 1e-5, (name, e)
    assert e["u3"] <= 1e-4, e



def test_fused_stage_unaligned_state_steps_bitwise_like_aligned():
    """A state 4 bytes off 16-byte alignment takes the fused stage's 4-byte
    assembly and stores; both forms sum the same terms in the same order, so
    two steps agree bitwise."""
    torch = _torch()
    from cadycore.lib import DryDycore

    g, num, Y32, D0 = _case("sphere")
    d = DryDycore(g, PAR, num, D0_col=D0)
    Ya = d.state(Y32)
    d.dss(Ya)
    buf = torch.zeros(d.S + 1, dtype=torch.float32, device=Ya.device)
    Yu = buf[1:]
    Yu.copy_(Ya)
    assert Ya.data_ptr() % 16 == 0 and Yu.data_ptr() % 16 != 0
    d.step(Ya, num.dt, 2)
    d.step(Yu, num.dt, 2)
    torch.cuda.synchronize()
    d.close()
    assert torch.equal(Ya.cpu(), Yu.cpu())


def test_deep_column_implicit_diffusion_takes_the_hook_kernel_stage():
    """nlev + 1 > 64 faces do not fit one wavefront: the default schedule is
    the thread-per-column stage, and it matches the oracle over 5 steps of a
    90-level column."""
    torch = _torch()
    from cadycore.lib import CA_TUNE_NEWTON

    g = G.columns(1, G.tanh_stretched_faces(90, 30000.0, 100.0), PAR.f_plane)
    num = P.Numerics(hyperdiff=False, dt=10.0, vert_diff=True, implicit_diffusion=True)
    Y32 = I.column_state(g, PAR, seed=0).astype(np.float32).astype(np.float64)
    d, o = _pair(g, num, np.array([40.0]))
    assert d.get_tuning(CA_TUNE_NEWTON) == 0
    Yo = Y32.copy()
    o.set_precomputed(Yo)
    o.step(Yo, num.dt, 5)
    Yg = d.state(Y32)
    d.step(Yg, num.dt, 5)
    torch.cuda.synchronize()
    e = errors(Yg.cpu().numpy().astype(np.float64), Yo, g)
    d.close()
    print("z90 column, 5 steps:", {k: f"{v:.2e}" for k, v in e.items()})
    for name in ("rho", "rhoe"):
        assert e[name] <= 1e-3, (name, e)
