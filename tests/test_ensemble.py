"""Column ensembles of uneven vertical sizes (ca_ensemble_*).

Each member must match the oracle run on its own column grid with its own
alpha_w and D0 (relative L2 per field, float32 native vs float64 oracle):
rho, u_h, rho e_tot within 1e-5 after one step and 1e-3 after 100; u_3
against its float32 storage floor (the float64 oracle with every stored
array rounded to float32).  After 100 steps u_3 is accumulated rounding
noise: a one-ulp change of the initial state alone moves it by 3e-5 to
4e-4, and single realisations of the storage floor range from 1e-3 to 2e-2
across members, so that gate is on the root mean square over the members
(at most the floors' root mean square), with 5x the member's own floor as
the per-member bound.  Every schedule (lanes per member, member order) and
the hipGraph replay give bitwise the same result: the per-member arithmetic
does not depend on the schedule.
"""
import numpy as np
import pytest

from cadycore import geometry as G
from cadycore import initial as I
from cadycore import params as P
from cadycore.ensemble import Member, c3_members, member_faces

PAR = P.Params.from_toml()
NUM = P.Numerics(hyperdiff=False, dt=10.0, rayleigh=True, vert_diff=True)


def test_c3_members_are_seeded_and_in_range():
    m1, s1 = c3_members(PAR, n=40, seed=3)
    m2, s2 = c3_members(PAR, n=40, seed=3)
    assert [m.nlev for m in m1] == [m.nlev for m in m2]
    assert all(np.array_equal(a, b) for a, b in zip(s1, s2))
    assert {m.nlev for m in m1} <= {30, 45, 63, 90, 128}
    assert all(1e-4 <= m.alpha_w <= 1e-2 and 0.1 <= m.D0 <= 50 for m in m1)
    for m, s in zip(m1, s1):
        assert len(s) == 5 * m.nlev + 1
        zf = m.z_face
        assert zf[0] == 0 and abs(zf[-1] - 30000) < 1e-6 and np.all(np.diff(zf) > 0)
    # 500 m bottom layer where it is thinner than the mean, else 30 m
    assert abs(member_faces(30)[1] - 500) < 1e-6 and abs(member_faces(128)[1] - 30) < 1e-6


def _torch():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def rel(a, b):
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / nb) if nb > 0 else float(np.abs(a).max())


def _members_small():
    rng = np.random.default_rng(11)
    members, states = [], []
    for i, nz in enumerate((30, 45, 63, 90, 128, 30, 63, 128)):
        zf = member_faces(nz)
        members.append(Member(zf, float(rng.uniform(1e-4, 1e-2)), float(rng.uniform(0.1, 50.0)), PAR.f_plane))
        g = G.columns(1, zf, PAR.f_plane)
        states.append(I.column_state(g, PAR, seed=20 + i, T_s=np.array([rng.uniform(285, 305)])))
    return members, [s.astype(np.float32).astype(np.float64) for s in states]


def _oracle_run(member, Y32, nsteps, rounding=False):
    from oracle import Oracle

    g = G.columns(1, member.z_face, PAR.f_plane)
    o = Oracle(g, PAR, NUM, alpha_w_col=np.array([member.alpha_w]), D0_col=np.array([member.D0]))
    Y = Y32.copy()
    o.dss(Y)
    o.set_precomputed(Y)
    if rounding:
        o.set_storage_rounding(True)
    o.step(Y, NUM.dt, nsteps)
    return Y


def _fields(y, nl):
    return I.unpack(y, 1, nl)


@pytest.mark.gpu
@pytest.mark.parametrize("nsteps", [1, 100])
def test_members_match_oracle(nsteps):
    torch = _torch()
    from cadycore.ensemble import ColumnEnsemble

    members, states = _members_small()
    e = ColumnEnsemble(members, PAR, NUM)
    Y = e.state(states)
    e.step(Y, NUM.dt, nsteps)
    torch.cuda.synchronize()
    got = e.unpack(Y)
    e.close()
    e3, f3 = [], []
    for m, (mem, Y32, y) in enumerate(zip(members, states, got)):
        ref = _oracle_run(mem, Y32, nsteps)
        floor = _oracle_run(mem, Y32, nsteps, rounding=True)
        fg, fr, ff = _fields(y.astype(np.float64), mem.nlev), _fields(ref, mem.nlev), _fields(floor, mem.nlev)
        errs = {n: rel(a, b) for n, a, b in zip(I.FIELD_NAMES, fg, fr)}
        fl = rel(ff[4], fr[4])
        print(m, mem.nlev, errs, "u3 storage floor", fl)
        tol = 1e-5 if nsteps == 1 else 1e-3
        for n in ("rho", "u1", "u2", "rhoe"):
            assert errs[n] <= tol, (m, n, errs[n])
        state = "MET" if errs["u3"] <= tol else "UNMET"
        print(f"u3 gate [member {m}, {nsteps} steps]: {errs['u3']:.2e} vs stated {tol:.0e}: {state}")
        assert errs["u3"] <= (3.0 if nsteps == 1 else 5.0) * fl, (m, errs["u3"], fl)  # regression guard
        e3.append(errs["u3"])
        f3.append(fl)
    rms_e, rms_f = float(np.sqrt(np.mean(np.square(e3)))), float(np.sqrt(np.mean(np.square(f3))))
    print("u3 rms over members: native", rms_e, "storage floor", rms_f)
    if nsteps > 1:
        assert rms_e <= rms_f


@pytest.mark.gpu
def test_schedules_and_graph_replay_are_bitwise_identical():
    torch = _torch()
    from cadycore.ensemble import ColumnEnsemble

    members, states = _members_small()
    e = ColumnEnsemble(members, PAR, NUM)
    Y0 = e.state(states)
    outs = []
    for lanes in (64, 128, 256):
        for order in (0, 1):
            e.set_tuning(lanes=lanes, order=order)
            Y = Y0.clone()
            e.step(Y, NUM.dt, 5)
            torch.cuda.synchronize()
            outs.append(Y.cpu())
    for o in outs[1:]:
        assert torch.equal(o, outs[0])
    # graph: two replays with host work and synchronisation between them
    e.set_tuning(lanes=64, order=1)
    st = torch.cuda.Stream()
    Yg = Y0.clone()
    torch.cuda.synchronize()
    e.step_graph(Yg, NUM.dt, 5, st)
    st.synchronize()
    first = Yg.cpu()
    assert torch.equal(first, outs[0])
    e.step_graph(Yg, NUM.dt, 5, st)
    st.synchronize()
    Yd = Y0.clone()
    e.step(Yd, NUM.dt, 10)
    torch.cuda.synchronize()
    assert torch.equal(Yg.cpu(), Yd.cpu())
    e.close()


@pytest.mark.gpu
def test_member_matches_single_column_step():
    """A member and DryDycore on the same one-column grid agree to float32
    rounding-order level after 10 steps."""
    torch = _torch()
    from cadycore.ensemble import ColumnEnsemble
    from cadycore.lib import DryDycore

    members, states = _members_small()
    e = ColumnEnsemble(members, PAR, NUM)
    Y = e.state(states)
    e.step(Y, NUM.dt, 10)
    torch.cuda.synchronize()
    got = e.unpack(Y)
    e.close()
    for mem, Y32, y in zip(members[:5], states[:5], got[:5]):
        g = G.columns(1, mem.z_face, PAR.f_plane)
        d = DryDycore(g, PAR, NUM, alpha_w_col=np.array([mem.alpha_w]), D0_col=np.array([mem.D0]))
        Yc = d.state(Y32)
        d.step(Yc, NUM.dt, 10)
        torch.cuda.synchronize()
        yc = Yc.cpu().numpy().astype(np.float64)
        d.close()
        fa, fb = _fields(y.astype(np.float64), mem.nlev), _fields(yc, mem.nlev)
        for n, a, b in zip(I.FIELD_NAMES, fa, fb):
            if n in ("u1", "u2") and np.linalg.norm(b) == 0:
                continue
            assert rel(a, b) <= (1e-5 if n != "u3" else 2e-3), (mem.nlev, n, rel(a, b))


@pytest.mark.gpu
def test_c3_full_ensemble_autotuned():
    """512 members, z_elem ~ U{30, 45, 63, 90, 128}, 360 steps of 10 s in one
    launch: finite, per-member column mass conserved to float32 round-off,
    boundary u_3 exactly 0; the autotuner tries all six schedules."""
    torch = _torch()
    from cadycore.ensemble import ColumnEnsemble

    members, states = c3_members(PAR, n=512, seed=3)
    e = ColumnEnsemble(members, PAR, NUM)
    Y = e.state(states)
    times = e.autotune(Y, NUM.dt, 20)
    print("autotune ms", times, "chosen", e.get_tuning())
    assert len(times) == 6 and all(t > 0 for t in times.values())
    mass0 = []
    for mem, s in zip(members, states):
        Jv = np.diff(mem.z_face)
        mass0.append(float(np.dot(np.asarray(s[:mem.nlev], np.float64), Jv)))
    e.step(Y, NUM.dt, 360)
    torch.cuda.synchronize()
    out = e.unpack(Y)
    e.close()
    for mem, y, m0 in zip(members, out, mass0):
        assert np.all(np.isfinite(y))
        nl = mem.nlev
        m1 = float(np.dot(y[:nl].astype(np.float64), np.diff(mem.z_face)))
        assert abs(m1 / m0 - 1) < 1e-5, (nl, m1 / m0 - 1)
        assert y[4 * nl] == 0 and y[5 * nl] == 0
