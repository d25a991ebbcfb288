"""ca_ensemble_adjoint_noalloc against ca_ensemble_adjoint, through the C ABI.

The no-allocation variant keeps its checkpoints and segment states in the
library's fixed per-GPU scratch array instead of handle allocations; it runs
the same kernels, so the cotangents must be bitwise identical.  Two calls in
a row, the second on a smaller ensemble (fewer members, fewer levels, fewer
steps), check that nothing of the first call's scratch contents leaks into
the second.  A request above the array's size must fail before any launch.
"""
import ctypes as C

import numpy as np
import pytest

from cadycore import geometry as G
from cadycore import initial as I
from cadycore import params as P
from cadycore.ensemble import ColumnEnsemble, Member, member_faces

PAR = P.Params.from_toml()
SCRATCH_FLOATS = 1 << 24


def _torch():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _ensemble(levels, seed):
    rng = np.random.default_rng(seed)
    members, states = [], []
    for i, nz in enumerate(levels):
        zf = member_faces(nz)
        members.append(Member(zf, float(rng.uniform(1e-4, 1e-2)), float(rng.uniform(0.1, 50.0)), PAR.f_plane))
        g = G.columns(1, zf, PAR.f_plane)
        states.append(I.column_state(g, PAR, seed=seed * 1000 + i, T_s=np.array([rng.uniform(285, 305)])))
    num = P.Numerics(hyperdiff=False, dt=10.0, rayleigh=True, vert_diff=True)
    return members, states, num


def _both(torch, levels, seed, nsteps, every):
    members, states, num = _ensemble(levels, seed)
    e = ColumnEnsemble(members, PAR, num)
    try:
        Y0 = e.state(states)
        rng = np.random.default_rng(seed + 1)
        YbN = torch.tensor(rng.standard_normal(e.S).astype(np.float32), device=e.device)
        ref = [x.cpu().numpy() for x in e.adjoint(Y0, YbN, num.dt, nsteps, every)]
        got = [x.cpu().numpy() for x in e.adjoint(Y0, YbN, num.dt, nsteps, every, noalloc=True)]
        torch.cuda.synchronize()
    finally:
        e.close()
    return ref, got


@pytest.mark.gpu
def test_noalloc_adjoint_matches_allocating_adjoint_on_two_calls():
    torch = _torch()
    # first call: 12 members up to 128 levels, 24 steps, checkpoints every 5
    ref1, got1 = _both(torch, (30, 45, 63, 90, 128, 30, 63, 128, 45, 90, 30, 63), 7, 24, 5)
    print("call 1: 12 members, 24 steps, every 5", flush=True)
    # second call: a smaller ensemble, fewer steps
    ref2, got2 = _both(torch, (30, 45, 30), 9, 10, 4)
    print("call 2: 3 members, 10 steps, every 4", flush=True)
    for ref, got in ((ref1, got1), (ref2, got2)):
        for name, a, b in zip(("Yb0", "D0b", "awb"), ref, got):
            assert np.all(np.isfinite(a)), name
            assert np.array_equal(a, b), f"{name}: max |diff| {np.max(np.abs(a - b)):.3e}"


@pytest.mark.gpu
def test_noalloc_adjoint_rejects_requests_above_the_scratch_size():
    torch = _torch()
    members, states, num = _ensemble((128, 128), 3)
    e = ColumnEnsemble(members, PAR, num)
    try:
        Y0 = e.state(states)
        YbN = torch.zeros_like(Y0)
        Yb0 = torch.full_like(Y0, 7.0)
        D0b = torch.full((2,), 7.0, device=e.device)
        awb = torch.full((2,), 7.0, device=e.device)
        nsteps = SCRATCH_FLOATS // e.S + 10  # every = nsteps: nsteps + 2 states
        st = e.lib.ca_ensemble_adjoint_noalloc(e.h, C.c_void_p(Y0.data_ptr()), C.c_void_p(YbN.data_ptr()), 10.0,
                                               nsteps, nsteps, C.c_void_p(Yb0.data_ptr()),
                                               C.c_void_p(D0b.data_ptr()), C.c_void_p(awb.data_ptr()), None)
        assert st == 1, st  # CA_ERR_ARG
        assert b"scratch" in e.lib.ca_ensemble_last_error(e.h)
        torch.cuda.synchronize()
        assert torch.all(Yb0 == 7.0) and torch.all(D0b == 7.0)  # nothing ran
    finally:
        e.close()
