"""Reverse mode of the sphere step through the C ABI (ca_dss_vjp,
ca_remaining_tendency_vjp, ca_step_adjoint) against float64 central
differences of the oracle.

Checks, on a small sphere (h2, 10 levels, hyperdiffusion, van Leer,
vertical diffusion with a D0 node field):
  * <DSS x, y> = <x, DSS^T y> (float32 round-off);
  * the pullback of remaining_tendency! along one random direction per
    prognostic field, and its D0 cotangent on a few columns;
  * the pullback of 4 steps (checkpoints every 2 and every 3, bitwise the
    same) along one random direction per field and for D0.
The directional derivatives are <Yb, v> of the native float32 pullback
against (F(Y + eps v) - F(Y - eps v)) / 2 eps of the float64 oracle with
eps = 1e-7: a random relative perturbation of rho or rho e_tot is strongly
acoustic, and at eps = 1e-5 it already moves u_3 enough to switch upwind
directions (the oracle's own difference then moves by ~5e-3).  The D0
cotangent of the steps is checked on the smooth functional sum_c r_c
(integral of T dz), as the workload uses; the gates allow for the float32
linearisation point.  The workload-sized check
(h8 z30, 96 steps, D0 gradient on 64 columns) compares with the committed
float64 differences in tests/golden/c4_fd_d0.npz (make_c4_fd.py).
"""
import os

import numpy as np
import pytest

from cadycore import geometry as G
from cadycore import initial as I
from cadycore import params as P

pytestmark = pytest.mark.gpu
PAR = P.Params.from_toml()
HERE = os.path.dirname(os.path.abspath(__file__))


def _torch():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _small(vert_diff=True):
    g = G.cubed_sphere(2, G.tanh_stretched_faces(10, 30000.0, 300.0), PAR.planet_radius, PAR.Omega)
    num = P.Numerics(dt=100.0, vert_diff=vert_diff)
    rng = np.random.default_rng(3)
    D0 = np.exp(rng.uniform(np.log(0.1), np.log(50.0), size=g.n_nodes))[g.node_id]
    Y0 = I.sphere_state(g, PAR, seed=2).astype(np.float32).astype(np.float64)
    return g, num, D0, Y0


def _fields(g):
    nC, nF = g.nlev * g.ncol, (g.nlev + 1) * g.ncol
    return [slice(i * nC, (i + 1) * nC) for i in range(4)] + [slice(4 * nC, 4 * nC + nF)]


def _direction(g, Y, field, seed):
    """A random direction in one field, scaled to that field (u_h: its rms;
    u_3: 1 m/s times the face spacing, zero on the boundary faces)."""
    rng = np.random.default_rng(seed)
    v = np.zeros_like(Y)
    sl = _fields(g)[field]
    r = rng.standard_normal(sl.stop - sl.start)
    if field in (0, 3):
        v[sl] = r * np.abs(Y[sl])
    elif field in (1, 2):  # covariant u_h: ~|u| times the element's metric length
        v[sl] = r * np.sqrt(np.mean(Y[sl] ** 2))
    else:
        _, _, Jvf = G.vertical_metrics(g.z_face)
        u = r.reshape(g.nlev + 1, g.ncol) * Jvf[:, None]
        u[0] = u[-1] = 0.0
        v[sl] = u.reshape(-1)
    return v


def _cotangent(g, ref, seed):
    """Random weights per field, scaled by 1 / rms of that field of ref; zero
    on the boundary faces (u_3 there is filtered to 0 and the remaining
    tendency hook writes 0; the oracle's hook fills those faces)."""
    rng = np.random.default_rng(seed)
    w = np.zeros_like(ref)
    for sl in _fields(g):
        s = np.sqrt(np.mean(ref[sl] ** 2))
        w[sl] = rng.standard_normal(sl.stop - sl.start) / (s if s > 0 else 1.0)
    nC = g.nlev * g.ncol
    w[4 * nC:4 * nC + g.ncol] = 0.0
    w[4 * nC + g.nlev * g.ncol:] = 0.0
    return w.astype(np.float32).astype(np.float64)


def test_dss_transpose_is_the_adjoint():
    torch = _torch()
    from cadycore.lib import DryDycore

    g, num, D0, _ = _small()
    d = DryDycore(g, PAR, num, D0_col=D0)
    gen = torch.Generator(device="cuda").manual_seed(1)
    x = torch.randn(d.S, device="cuda", generator=gen)
    y = torch.randn(d.S, device="cuda", generator=gen)
    Dx, Dty = x.clone(), y.clone()
    d.dss(Dx)
    d.dss_vjp(Dty)
    torch.cuda.synchronize()
    lhs = float(torch.dot(Dx.double(), y.double()))
    rhs = float(torch.dot(x.double(), Dty.double()))
    scale = float(Dx.double().norm() * y.double().norm())
    print("<DSS x, y>", lhs, "<x, DSS^T y>", rhs, "rel", abs(lhs - rhs) / scale)
    assert abs(lhs - rhs) <= 1e-6 * scale
    d.close()


def _oracle_T(o, Y):
    o.set_precomputed(Y)
    return o.remaining_tendency(Y)


@pytest.mark.parametrize("field", [0, 1, 2, 3, 4])
def test_remaining_tendency_pullback_directional(field):
    torch = _torch()
    from cadycore.lib import DryDycore
    from oracle import Oracle

    g, num, D0, Y0 = _small()
    o = Oracle(g, PAR, num, D0_col=D0)
    Y = Y0.copy()
    o.dss(Y)
    Y = Y.astype(np.float32).astype(np.float64)
    T0 = _oracle_T(o, Y)
    w = _cotangent(g, T0, 11)
    d = DryDycore(g, PAR, num, D0_col=D0)
    Yb = d.zeros()
    d.remaining_tendency_vjp(Yb, d.state(w), d.state(Y))
    torch.cuda.synchronize()
    yb = Yb.cpu().numpy().astype(np.float64)
    v = _direction(g, Y, field, 20 + field)
    eps = 1e-7
    fd = (w @ _oracle_T(o, Y + eps * v) - w @ _oracle_T(o, Y - eps * v)) / (2 * eps)
    ad = float(yb @ v)
    rel = abs(ad - fd) / abs(fd)
    print(f"field {field}: <Yb, v> {ad:.8e} fd {fd:.8e} rel {rel:.2e}")
    d.close()
    assert rel <= 1e-3


def test_remaining_tendency_pullback_d0():
    torch = _torch()
    from cadycore.lib import DryDycore
    from oracle import Oracle

    g, num, D0, Y0 = _small()
    Y = Y0.copy()
    Oracle(g, PAR, num, D0_col=D0).dss(Y)
    Y = Y.astype(np.float32).astype(np.float64)
    T0 = _oracle_T(Oracle(g, PAR, num, D0_col=D0), Y)
    w = _cotangent(g, T0, 12)
    d = DryDycore(g, PAR, num, D0_col=D0)
    Yb = d.zeros()
    D0b = torch.zeros(g.ncol, dtype=torch.float32, device="cuda")
    d.remaining_tendency_vjp(Yb, d.state(w), d.state(Y), D0b)
    torch.cuda.synchronize()
    gb = D0b.cpu().numpy().astype(np.float64)
    cols = np.random.default_rng(5).choice(g.ncol, size=6, replace=False)
    for c in cols:
        h = 1e-4 * D0[c]
        Dp, Dm = D0.copy(), D0.copy()
        Dp[c] += h
        Dm[c] -= h
        fd = (w @ _oracle_T(Oracle(g, PAR, num, D0_col=Dp), Y) - w @ _oracle_T(Oracle(g, PAR, num, D0_col=Dm), Y)) / (2 * h)
        rel = abs(gb[c] - fd) / max(abs(fd), 1e-300)
        print(f"col {c}: native {gb[c]:.8e} fd {fd:.8e} rel {rel:.2e}")
        assert rel <= 1e-3
    d.close()


def _oracle_steps(g, num, D0, Y0, nsteps):
    from oracle import Oracle

    o = Oracle(g, PAR, num, D0_col=D0)
    Y = Y0.copy()
    o.dss(Y)
    o.set_precomputed(Y)
    o.step(Y, num.dt, nsteps)
    return Y


NSTEPS = 4


@pytest.fixture(scope="module")
def step_pullback():
    torch = _torch()
    from cadycore.lib import DryDycore

    g, num, D0, Y0 = _small()
    YN = _oracle_steps(g, num, D0, Y0, NSTEPS)
    w = _cotangent(g, YN, 13)
    d = DryDycore(g, PAR, num, D0_col=D0)
    outs = []
    for every in (2, 3, 0):  # 0: the forward keeps every step's records, no recomputation
        D0b = torch.zeros(g.ncol, dtype=torch.float32, device="cuda")
        Yb0 = d.step_adjoint(d.state(Y0), d.state(w), num.dt, NSTEPS, every=every, D0b=D0b)
        torch.cuda.synchronize()
        outs.append((Yb0.cpu().numpy(), D0b.cpu().numpy()))
    d.close()
    return g, num, D0, Y0, w, outs


def test_step_pullback_checkpoint_spacing_bitwise(step_pullback):
    """Checkpoints every 2 or 3 steps, or all stage records kept (every = 0):
    the same records feed the same reverse steps, so the cotangents agree
    bitwise."""
    *_, outs = step_pullback
    for o in outs[1:]:
        assert np.array_equal(outs[0][0], o[0]) and np.array_equal(outs[0][1], o[1])
    assert np.all(np.isfinite(outs[0][0]))


@pytest.mark.parametrize("field", [0, 1, 2, 3, 4])
def test_step_pullback_directional(step_pullback, field):
    g, num, D0, Y0, w, outs = step_pullback
    yb = outs[0][0].astype(np.float64)
    v = _direction(g, Y0, field, 30 + field)
    eps = 1e-7
    fd = (w @ _oracle_steps(g, num, D0, Y0 + eps * v, NSTEPS) - w @ _oracle_steps(g, num, D0, Y0 - eps * v, NSTEPS)) / (2 * eps)
    ad = float(yb @ v)
    rel = abs(ad - fd) / abs(fd)
    print(f"field {field}: <Yb0, v> {ad:.8e} fd {fd:.8e} rel {rel:.2e}")
    assert rel <= 1e-3


def test_step_pullback_d0():
    """D0 cotangent of sum_c r_c (integral of T dz) after 4 steps, column by
    column, against float64 central differences (h = 1e-3 D0)."""
    torch = _torch()
    from cadycore.lib import DryDycore
    from cadycore.workloads import column_integral

    g, num, D0, Y0 = _small()
    r = np.random.default_rng(7).choice([-1.0, 1.0], size=g.ncol) * np.random.default_rng(8).uniform(0.5, 2.0, g.ncol)
    d = DryDycore(g, PAR, num, D0_col=D0)
    YN = d.state(Y0)
    d.step(YN, num.dt, NSTEPS)
    Yd = YN.double().requires_grad_(True)
    L = (torch.tensor(r, dtype=torch.float64, device="cuda") * column_integral(Yd, g, PAR, xp=torch)).sum()
    L.backward()
    D0b = torch.zeros(g.ncol, dtype=torch.float32, device="cuda")
    d.step_adjoint(d.state(Y0), Yd.grad.float().contiguous(), num.dt, NSTEPS, every=2, D0b=D0b)
    torch.cuda.synchronize()
    gb = D0b.cpu().numpy().astype(np.float64)
    d.close()

    def F(D):
        return float(r @ column_integral(_oracle_steps(g, num, D, Y0, NSTEPS), g, PAR))

    cols = np.random.default_rng(6).choice(g.ncol, size=6, replace=False)
    for c in cols:
        h = 1e-3 * D0[c]
        Dp, Dm = D0.copy(), D0.copy()
        Dp[c] += h
        Dm[c] -= h
        fd = (F(Dp) - F(Dm)) / (2 * h)
        rel = abs(gb[c] - fd) / max(abs(fd), 1e-300)
        print(f"col {c}: native {gb[c]:.8e} fd {fd:.8e} rel {rel:.2e}")
        assert rel <= 1e-4


def test_c4_d0_gradient_matches_float64_fixture():
    """The differentiation workload (h8 z30, 96 steps of dt = 150 s,
    checkpoints every 8, squared column-integrated temperature loss) through
    the autograd path (StepSolve: ca_step_record / ca_step_pullback): the D0
    gradient on the fixture's columns against its float64 central
    differences (tests/golden/make_c4_fd.py), relative L2 <= 1e-4 over the
    columns and <= 1e-3 for each."""
    torch = _torch()
    path = os.path.join(HERE, "golden", "c4_fd_d0.npz")
    if not os.path.exists(path):
        pytest.skip("fixture tests/golden/c4_fd_d0.npz not generated")
    from cadycore.adjoint import StepSolve
    from cadycore.lib import DryDycore
    from cadycore.workloads import c4_case, c4_loss

    z = np.load(path)
    n = int(z["done"])
    cols, fd = z["cols"][:n], z["fd"][:n]
    case = c4_case(PAR)
    assert np.allclose(case["D0"][cols], z["D0"][:n])
    d = DryDycore(case["grid"], PAR, case["numerics"], D0_col=case["D0"])
    D0 = torch.tensor(case["D0"], dtype=torch.float64, requires_grad=True)
    Y = StepSolve.apply(d.state(case["Y0"]), D0, d, case["dt"], case["nsteps"], case["every"])
    L = c4_loss(Y.double(), case, xp=torch)
    L.backward()
    torch.cuda.synchronize()
    g = D0.grad.numpy()[cols]
    d.close()
    rel = np.abs(g - fd) / np.abs(fd)
    vec = np.linalg.norm(g - fd) / np.linalg.norm(fd)
    print(f"loss {float(L):.8e} (float64 oracle {float(z['L0']):.8e}); {n} columns: relative L2 {vec:.2e}, "
          f"max {rel.max():.2e}, median {np.median(rel):.2e}")
    assert abs(float(L) / float(z["L0"]) - 1) < 1e-5
    assert vec <= 1e-4 and rel.max() <= 1e-3


@pytest.mark.parametrize("nlev,newton", [(40, 1), (40, 0), (80, 0)])
def test_step_pullback_deep_columns_and_schedules(nlev, newton):
    """The pullback on the other reverse Newton-stage shapes (one column per
    64-lane wavefront at 33..64 levels, 128 lanes at 65..128) and with the
    stage recorded by the thread-per-column forward (CA_TUNE_NEWTON 0),
    against float64 central differences of the oracle."""
    torch = _torch()
    from cadycore.lib import CA_TUNE_NEWTON, DryDycore

    g = G.cubed_sphere(2, G.tanh_stretched_faces(nlev, 30000.0, 100.0), PAR.planet_radius, PAR.Omega)
    num = P.Numerics(dt=100.0, vert_diff=True)
    rng = np.random.default_rng(3)
    D0 = np.exp(rng.uniform(np.log(0.1), np.log(50.0), size=g.n_nodes))[g.node_id]
    Y0 = I.sphere_state(g, PAR, seed=2).astype(np.float32).astype(np.float64)
    nsteps = 3
    w = _cotangent(g, _oracle_steps(g, num, D0, Y0, nsteps), 17)
    d = DryDycore(g, PAR, num, D0_col=D0)
    d.set_tuning(CA_TUNE_NEWTON, newton)
    yb = d.step_adjoint(d.state(Y0), d.state(w), num.dt, nsteps, every=2).cpu().numpy().astype(np.float64)
    d.close()
    assert np.all(np.isfinite(yb))
    for field in (3, 4):
        v = _direction(g, Y0, field, 40 + field)
        eps = 1e-7
        fd = (w @ _oracle_steps(g, num, D0, Y0 + eps * v, nsteps)
              - w @ _oracle_steps(g, num, D0, Y0 - eps * v, nsteps)) / (2 * eps)
        rel = abs(float(yb @ v) - fd) / abs(fd)
        print(f"nlev {nlev} newton {newton} field {field}: rel {rel:.2e}")
        assert rel <= 1e-3
