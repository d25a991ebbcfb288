"""Reverse mode of the column-ensemble step (ca_ensemble_adjoint) against
float64 central finite differences of the oracle.

Workload: columns of 30 levels over 30 km (dz_bottom 500 m), dt = 150 s,
96 steps with a checkpoint every 8, van Leer upwinding, vertical diffusion
with a per-column D0 drawn log-uniform in [0.1, 50] m^2/s.  The loss is the
mean over columns of (integral of T dz - target)^2 with a seeded target.
The native gradient differentiates the float32 step; the finite
differences run the float64 oracle, so the comparison also bounds the
float32 effects.  For that squared loss the gate is on the gradient vector
(relative L2 over the columns <= 1e-4): a column's own factor
(integral - target) is ~2e3 against an integral of ~7.5e6, so the float32
trajectory moves it, and that column's gradient, by ~1e-3 relative.  A
linear loss (sum of w_m times the integral) has no such factor and pins the
reverse mode column by column (<= 1e-4 each).
"""
import numpy as np
import pytest

from cadycore import geometry as G
from cadycore import initial as I
from cadycore import params as P
from cadycore.ensemble import Member, member_faces

PAR = P.Params.from_toml()
DT, NSTEPS, EVERY = 150.0, 96, 8


def _torch():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _setup(n=6, seed=4, rayleigh=False, w_amp=0.0):
    """n columns with seeded D0, alpha_w, surface temperature and target;
    w_amp (m/s) adds a vertical velocity sin(pi z / z_top) on the faces."""
    rng = np.random.default_rng(seed)
    zf = member_faces(30)
    K = np.exp(rng.uniform(np.log(0.1), np.log(50.0), size=n))
    aw = rng.uniform(1e-4, 1e-2, size=n)
    members = [Member(zf, float(aw[i]), float(K[i]), PAR.f_plane) for i in range(n)]
    states = []
    for i in range(n):
        g = G.columns(1, zf, PAR.f_plane)
        states.append(I.column_state(g, PAR, seed=100 + i, T_s=np.array([rng.uniform(285, 305)])))
    if w_amp:
        nl = len(zf) - 1
        zc = 0.5 * (zf[1:] + zf[:-1])
        jvf = zc[1:] - zc[:-1]
        for s in states:
            s[4 * nl + 1:5 * nl] += w_amp * np.sin(np.pi * zf[1:-1] / zf[-1]) * jvf
    states = [s.astype(np.float32).astype(np.float64) for s in states]
    num = P.Numerics(hyperdiff=False, dt=DT, rayleigh=rayleigh, vert_diff=True)
    dz = np.diff(zf)
    integ = [float(np.sum(_T_np(s, zf) * dz)) for s in states]
    target = np.array(integ) + rng.normal(0.0, 2000.0, size=n)
    return members, states, num, target


def _T_np(y, zf):
    nl = len(zf) - 1
    rho, u1, u2, rhoe, u3 = y[:nl], y[nl:2 * nl], y[2 * nl:3 * nl], y[3 * nl:4 * nl], y[4 * nl:]
    zc = 0.5 * (zf[1:] + zf[:-1])
    jvf = np.empty(nl + 1)
    jvf[1:-1] = zc[1:] - zc[:-1]
    jvf[0], jvf[-1] = zf[1] - zf[0], zf[-1] - zf[-2]
    u3i = u3.copy()
    u3i[0] = u3i[-1] = 0.0
    K = 0.5 * (u1 ** 2 + u2 ** 2) + 0.25 * (u3i[:-1] ** 2 / jvf[:-1] ** 2 + u3i[1:] ** 2 / jvf[1:] ** 2)
    T = PAR.T_0 + (rhoe / rho - K - PAR.grav * zc + PAR.R_d * PAR.T_0) / PAR.cv_d
    return np.maximum(T, PAR.T_min_sgs)


def _oracle_final(member, y0, num, D0=None, aw=None, dy=None, nsteps=NSTEPS):
    """state after nsteps float64 oracle steps."""
    from oracle import Oracle

    g = G.columns(1, member.z_face, PAR.f_plane)
    o = Oracle(g, PAR, num, alpha_w_col=np.array([member.alpha_w if aw is None else aw]),
               D0_col=np.array([member.D0 if D0 is None else D0]))
    Y = y0.copy() if dy is None else y0 + dy
    o.dss(Y)
    o.set_precomputed(Y)
    o.step(Y, DT, nsteps)
    return Y


def _oracle_integral(member, y0, num, **kw):
    """integral of T dz after NSTEPS float64 oracle steps."""
    Y = _oracle_final(member, y0, num, **kw)
    return float(np.sum(_T_np(Y, member.z_face) * np.diff(member.z_face)))


def _oracle_w2(member, y0, num, **kw):
    """sum of u_3^2 over the faces after NSTEPS float64 oracle steps."""
    Y = _oracle_final(member, y0, num, **kw)
    return float(np.sum(Y[4 * member.nlev:] ** 2))


def _oracle_loss_term(member, y0, num, target, **kw):
    """(integral T dz - target)^2 after NSTEPS float64 oracle steps."""
    return (_oracle_integral(member, y0, num, **kw) - target) ** 2


def _native_grads(members, states, num, target, weights=None, functional="integral", nsteps=NSTEPS):
    """Gradients of the squared loss, or with ``weights`` of the loss
    sum_m w_m F_m (F = integral of T dz, or the sum of u_3^2), through the
    native reverse mode."""
    torch = _torch()
    from cadycore.adjoint import EnsembleSolve, column_loss, column_temperature
    from cadycore.ensemble import ColumnEnsemble

    e = ColumnEnsemble(members, PAR, num)
    Y0 = e.state(states).requires_grad_(True)
    D0 = torch.tensor([m.D0 for m in members], dtype=torch.float64, requires_grad=True)
    aw = torch.tensor([m.alpha_w for m in members], dtype=torch.float64, requires_grad=True)
    Y = EnsembleSolve.apply(Y0, D0, aw, e, DT, nsteps, EVERY)
    if weights is None:
        L = column_loss(Y.double(), members, PAR, target)
    elif functional == "integral":
        Ts = column_temperature(Y.double(), members, PAR)
        L = sum(float(w) * torch.sum(T * torch.tensor(np.diff(m.z_face), dtype=T.dtype, device=T.device))
                for w, T, m in zip(weights, Ts, members))
    else:
        Yd, L, off = Y.double(), 0.0, 0
        for w, m in zip(weights, members):
            nl = m.nlev
            L = L + float(w) * torch.sum(Yd[off + 4 * nl:off + 5 * nl + 1] ** 2)
            off += 5 * nl + 1
    L.backward()
    torch.cuda.synchronize()
    out = (float(L), D0.grad.numpy().copy(), aw.grad.numpy().copy(), Y0.grad.detach().cpu().numpy().copy())
    e.close()
    return out


@pytest.mark.gpu
def test_d0_gradient_matches_float64_central_differences():
    members, states, num, target = _setup()
    n = len(members)
    L, gD0, _, _ = _native_grads(members, states, num, target)
    fds = []
    for i, (m, y0) in enumerate(zip(members, states)):
        h = 1e-3 * m.D0
        lp = _oracle_loss_term(m, y0, num, target[i], D0=m.D0 + h)
        lm = _oracle_loss_term(m, y0, num, target[i], D0=m.D0 - h)
        fds.append((lp - lm) / (2 * h) / n)
        print(i, "D0", m.D0, "native", gD0[i], "fd", fds[-1])
    fds = np.array(fds)
    rel = np.linalg.norm(gD0 - fds) / np.linalg.norm(fds)
    print("gradient relative L2 error", rel)
    assert rel <= 1e-4


@pytest.mark.gpu
@pytest.mark.parametrize("param", ["D0", "alpha_w"])
def test_weighted_loss_gradients_per_column(param):
    """Loss sum_m w_m F_m, column by column within 1e-4 of the float64
    central difference: the D0 gradient of F = integral of T dz over 96
    steps, and (sponge on) the alpha_w gradient of F = sum of u_3^2 over 2
    steps from a vertical velocity of 1 m/s amplitude, which the sponge
    damps.  (The integral of T barely depends on alpha_w, ~1e-12 relative
    per unit relative change, too flat for a float64 difference.  Later in
    the run the sponge has removed most of u_3, sum u_3^2 falls by 1e5 over
    8 steps, and the float64 differences of the residual stop converging.)"""
    rayleigh = param == "alpha_w"
    members, states, num, target = _setup(n=4, seed=7 if rayleigh else 13, rayleigh=rayleigh,
                                          w_amp=1.0 if rayleigh else 0.0)
    w = np.random.default_rng(2).choice([-1.0, 1.0], size=len(members)) * np.random.default_rng(3).uniform(0.5, 2, 4)
    functional = "w2" if rayleigh else "integral"
    ns = 2 if rayleigh else NSTEPS
    _, gD0, gaw, _ = _native_grads(members, states, num, target, weights=w, functional=functional, nsteps=ns)
    g = gaw if rayleigh else gD0
    F = _oracle_w2 if rayleigh else _oracle_integral
    for i, (m, y0) in enumerate(zip(members, states)):
        p0 = m.alpha_w if rayleigh else m.D0
        h = 1e-3 * p0
        kw = (lambda v: {"aw": v, "nsteps": ns}) if rayleigh else (lambda v: {"D0": v})
        fd = w[i] * (F(m, y0, num, **kw(p0 + h)) - F(m, y0, num, **kw(p0 - h))) / (2 * h)
        rel = abs(g[i] - fd) / max(abs(fd), 1e-300)
        print(i, param, p0, "native", g[i], "fd", fd, "rel", rel)
        assert rel <= 1e-4, (i, rel)


@pytest.mark.gpu
def test_initial_state_cotangent_matches_directional_difference():
    members, states, num, target = _setup(n=3, seed=9)
    n = len(members)
    _, _, _, gY0 = _native_grads(members, states, num, target)
    rng = np.random.default_rng(5)
    off = 0
    for i, (m, y0) in enumerate(zip(members, states)):
        nl = m.nlev
        sz = 5 * nl + 1
        g = gY0[off:off + sz].astype(np.float64)
        off += sz
        # a smooth perturbation of rho e_tot and rho (relative 1e-6), u_h
        v = np.zeros(sz)
        v[3 * nl:4 * nl] = y0[3 * nl:4 * nl] * rng.standard_normal(nl)
        v[:nl] = y0[:nl] * rng.standard_normal(nl)
        v[nl:2 * nl] = rng.standard_normal(nl)
        eps = 1e-6
        lp = _oracle_loss_term(m, y0, num, target[i], dy=eps * v)
        lm = _oracle_loss_term(m, y0, num, target[i], dy=-eps * v)
        fd = (lp - lm) / (2 * eps) / n
        ad = float(np.dot(g, v))
        rel = abs(ad - fd) / max(abs(fd), 1e-300)
        print(i, "<gY0, v>", ad, "fd", fd, "rel", rel)
        assert rel <= 1e-3


@pytest.mark.gpu
def test_checkpoint_spacing_does_not_change_the_gradient():
    torch = _torch()
    from cadycore.ensemble import ColumnEnsemble

    members, states, num, target = _setup(n=3, seed=11)
    e = ColumnEnsemble(members, PAR, num)
    Y0 = e.state(states)
    YbN = torch.randn(e.S, device=Y0.device, dtype=torch.float32)
    outs = [e.adjoint(Y0, YbN, DT, 20, every) for every in (1, 4, 7, 20)]
    torch.cuda.synchronize()
    for o in outs[1:]:
        for a, b in zip(o, outs[0]):
            assert torch.equal(a, b)
    e.close()
