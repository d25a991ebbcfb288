"""Workload definitions shared by the benchmark, the tests and the fixture
scripts: the differentiation workload (global h8 z30 sphere, per-column
vertical diffusivity, column-integrated temperature loss) and its loss.

The loss restates the calibration objective of the workload: the mean over
horizontal nodes of (integral of T dz - target)^2, T from the dry equation of
state of the final state (precomputed_quantities.jl:786-796 with
e_int = e_tot - K - Phi).  Shared element-edge nodes are stored in several
columns; each copy carries weight 1 / (copies x nodes), so every node counts
once.
"""
from __future__ import annotations

import numpy as np

from . import geometry as G
from . import initial as I
from . import params as P


def column_temperature(Y, grid, params, xp=np):
    """T [nlev, ncol] of a packed sphere state (numpy or torch via ``xp``):
    K = u_h . g^{-1} u_h / (2 s^2) + the centre mean of u_3^2 g^33 / 2."""
    nl, nc = grid.nlev, grid.ncol
    n = nl * nc
    rho = Y[0:n].reshape(nl, nc)
    u1 = Y[n:2 * n].reshape(nl, nc)
    u2 = Y[2 * n:3 * n].reshape(nl, nc)
    rhoe = Y[3 * n:4 * n].reshape(nl, nc)
    u3 = Y[4 * n:].reshape(nl + 1, nc)
    zc, _, Jvf = G.vertical_metrics(grid.z_face)
    R = grid.radius
    s = (R + zc) / R if R > 0 else np.ones_like(zc)
    is2 = 1.0 / s ** 2

    def arr(a):
        return xp.asarray(a) if xp is np else xp.tensor(a, dtype=Y.dtype, device=Y.device)

    g11, g12, g22 = (arr(grid.gup[t][None, :]) for t in range(3))
    is2 = arr(is2[:, None])
    ct1 = is2 * (g11 * u1 + g12 * u2)
    ct2 = is2 * (g12 * u1 + g22 * u2)
    g33 = arr((1.0 / Jvf ** 2)[:, None])
    mask = np.ones((nl + 1, 1))
    mask[0] = mask[-1] = 0.0
    u3i = u3 * arr(mask)
    Kv = 0.25 * (u3i[:-1] ** 2 * g33[:-1] + u3i[1:] ** 2 * g33[1:])
    K = 0.5 * (u1 * ct1 + u2 * ct2) + Kv
    phi = arr((params.grav * zc)[:, None])
    T = params.T_0 + (rhoe / rho - K - phi + params.R_d * params.T_0) / params.cv_d
    return np.maximum(T, params.T_min_sgs) if xp is np else T.clamp(min=params.T_min_sgs)


def column_integral(Y, grid, params, xp=np):
    """integral of T dz per column."""
    dz = np.diff(grid.z_face)
    T = column_temperature(Y, grid, params, xp)
    w = xp.asarray(dz[:, None]) if xp is np else xp.tensor(dz[:, None], dtype=Y.dtype, device=Y.device)
    return (T * w).sum(0)


def c4_loss(Y, case, xp=np):
    I_ = column_integral(Y, case["grid"], case["params"], xp)
    t = case["target"]
    w = case["weights"]
    if xp is not np:
        t = xp.tensor(t, dtype=Y.dtype, device=Y.device)
        w = xp.tensor(w, dtype=Y.dtype, device=Y.device)
    return (w * (I_ - t) ** 2).sum()


def c4_case(params: P.Params, seed: int = 4):
    """The differentiation workload: dry global sphere h_elem = 8, Nq = 4,
    z_elem = 30 over 30 km (tanh stretching, dz_bottom = 500 m), the seeded
    balanced state with the zonal jet and 1 K noise (seed 4), hyperdiffusion
    and van Leer upwinding as in the parity profile, vertical diffusion with
    D0 per horizontal node log-uniform in [0.1, 50] m^2/s, dt = 150 s, 96
    steps, checkpoints every 8; target = the initial column integral of T
    plus N(0, 2000 K m) per node."""
    zf = G.tanh_stretched_faces(30, 30000.0, 500.0)
    grid = G.cubed_sphere(8, zf, params.planet_radius, params.Omega)
    num = P.Numerics(dt=150.0, vert_diff=True)
    rng = np.random.default_rng(seed)
    D0_node = np.exp(rng.uniform(np.log(0.1), np.log(50.0), size=grid.n_nodes))
    Y0 = I.sphere_state(grid, params, seed=seed).astype(np.float32).astype(np.float64)
    mult = np.bincount(grid.node_id, minlength=grid.n_nodes)[grid.node_id].astype(np.float64)
    I0 = column_integral(Y0, grid, params)
    noise = rng.normal(0.0, 2000.0, size=grid.n_nodes)
    # the target is a node field: the copies of a node share its value
    I0_node = np.zeros(grid.n_nodes)
    np.add.at(I0_node, grid.node_id, I0 / mult)
    target = (I0_node + noise)[grid.node_id]
    return dict(grid=grid, numerics=num, params=params, Y0=Y0, D0=D0_node[grid.node_id], target=target,
                weights=1.0 / (mult * grid.n_nodes), mult=mult, dt=150.0, nsteps=96, every=8)
