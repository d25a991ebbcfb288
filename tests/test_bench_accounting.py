"""bench.py's byte accounting and the ensemble workload's CPU baseline
(host only): every launched kernel class has a byte count, the per-step sum
is the compulsory traffic of the step's pipeline, and the ensemble's CPU leg
steps the oracle columns it is given."""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "climaatmos.jl_amd"))

import bench  # noqa: E402


def _sizes(h_elem=30, nlev=63):
    ncol = 6 * h_elem * h_elem * 16
    nC, nF = nlev * ncol, (nlev + 1) * ncol
    return 4 * nC + nF, nC, nF


def test_every_launched_kernel_has_bytes():
    S, nC, nF = _sizes()
    kb = bench.kernel_bytes(S, nC, nF)
    assert set(kb) == set(bench.LAUNCHES_PER_STEP)
    assert all(v > 0 for v in kb.values())


def test_step_traffic_in_state_sizes():
    # stages 1-3: Laplacians (S + 4nC read, 4nC written), their DSS (8nC), the
    # tendency (2S + 4nC) and its DSS (2S); stage 4: Laplacians, their DSS and
    # the fused tendency + final combination (5S + 4nC); Newton stages
    # (2S + S+fi), (3S+fi + S+fi), (4S+2fi + S); the state DSS (2S)
    S, nC, nF = _sizes()
    fi = 2 * nC + nF
    want = (4 * ((S + 4 * nC) + 8 * nC) + 3 * ((2 * S + 4 * nC) + 2 * S) + (5 * S + 4 * nC)
            + (2 * S + S + fi) + (3 * S + fi + S + fi) + (4 * S + 2 * fi + S) + 2 * S)
    got = sum(bench.LAUNCHES_PER_STEP[k] * v for k, v in bench.kernel_bytes(S, nC, nF).items()) / 4
    assert abs(got - want) < 1e-6 * want
    assert abs(got / S - 50.77) < 0.01


def test_ensemble_cpu_baseline_steps_the_members():
    from cadycore.ensemble import c3_members
    from cadycore.params import Params

    par = Params.from_toml()
    members, states = c3_members(par, n=4, seed=3)
    out = bench.cpu_baseline_c3(par, members, states, 10.0, 2)
    assert out["kind"] == "port" and out["unit"] == "ensemble steps/s"
    assert out["value"] > 0 and np.isfinite(out["column_steps_per_s"])
