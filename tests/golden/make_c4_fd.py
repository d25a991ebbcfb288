"""Float64 central finite differences of the differentiation workload's loss
(cadycore.workloads.c4_case: global h8 z30 sphere, 96 steps of dt = 150 s)
with respect to D0 at 64 seeded interior columns, through the oracle.

Writes tests/golden/c4_fd_d0.npz: the columns, their D0, the relative step,
the base loss and the differences.  A node interior to an element is stored
in one column, so perturbing that column's D0 keeps the field a node field.
Runs ~2 oracle runs per column (about 1 CPU-minute each on 8 cores); the file
is rewritten after every column, so an interrupted run keeps what it has.

    python tests/golden/make_c4_fd.py [ncols]
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "climaatmos.jl_amd"), os.path.join(ROOT, "oracle")]

from cadycore import params as P  # noqa: E402
from cadycore.workloads import c4_case, c4_loss  # noqa: E402
from oracle import Oracle  # noqa: E402

OUT = os.path.join(HERE, "c4_fd_d0.npz")
REL_H = 1e-3


def loss(case, D0):
    o = Oracle(case["grid"], case["params"], case["numerics"], D0_col=D0)
    Y = case["Y0"].copy()
    o.dss(Y)
    o.set_precomputed(Y)
    o.step(Y, case["dt"], case["nsteps"])
    return float(c4_loss(Y, case))


def main(ncols=64):
    par = P.Params.from_toml()
    case = c4_case(par)
    g = case["grid"]
    interior = np.flatnonzero(case["mult"] == 1)
    cols = np.sort(np.random.default_rng(40).choice(interior, size=ncols, replace=False))
    D0 = case["D0"]
    t0 = time.time()
    L0 = loss(case, D0)
    print(f"base loss {L0:.10e} ({time.time() - t0:.0f} s)", flush=True)
    fd = np.full(ncols, np.nan)
    for i, c in enumerate(cols):
        h = REL_H * D0[c]
        Dp, Dm = D0.copy(), D0.copy()
        Dp[c] += h
        Dm[c] -= h
        fd[i] = (loss(case, Dp) - loss(case, Dm)) / (2 * h)
        np.savez(OUT, cols=cols, D0=D0[cols], fd=fd, rel_h=REL_H, L0=L0, done=i + 1,
                 h_elem=g.h_elem, nlev=g.nlev, dt=case["dt"], nsteps=case["nsteps"])
        print(f"{i + 1}/{ncols} col {c} D0 {D0[c]:.4g} dL/dD0 {fd[i]:.6e} ({time.time() - t0:.0f} s)", flush=True)


if __name__ == "__main__":
    main(int(sys.argv[1]) if len(sys.argv) > 1 else 64)
