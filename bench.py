#!/usr/bin/env python3
"""Benchmark of the native dry-dycore step on one MI355X.

Workload (the headline configuration): dry global cubed sphere, h_elem = 30,
Nq = 4, z_elem = 63 stretched to 30 km (dz_bottom = 30 m), Float32, dt = 60 s,
balanced profile + zonal jet + seeded 1 K temperature noise below 5 km.
A step is one ARS343 IMEX step (4 explicit tendencies with hyperdiffusion and
DSS, 3 implicit Newton stages, upwind correction) on the whole sphere.

Prints ONE JSON line.  ``value`` is simulated years per wall-clock day (SYPD)
of the timed steps with the state resident in HBM.  ``roofline`` is for the
kernel that takes the most device time, from HIP events recorded on the
stream it is launched on.  ``cpu_baseline`` times the CPU restatement
(``oracle/``) on the host on a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "climaatmos.jl_amd"))

METRIC = "SYPD + achieved HBM GB/s, dry global h30 z63 Nq4 FP32; adjoint/forward ratio"
HBM_PEAK_GBPS = 8000.0
# launches per ARS343 step of each kernel of kernel_bytes (the step's
# compulsory traffic is their weighted sum: 51.4 state sizes at h30 z63).
# The tendencies of stages 1-3 are DSSed; the last explicit launch writes the
# step's final combination, which the end-of-step DSS of the state averages.
LAUNCHES_PER_STEP = {"k_laplacians": 4, "k_dss(laplacians)": 4, "k_explicit": 3, "k_explicit_final": 1,
                     "k_dss(tendency)": 3, "k_newton_stage": 3, "k_dss(state)": 1}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--h-elem", type=int, default=30)
    ap.add_argument("--z-elem", type=int, default=63)
    ap.add_argument("--dt", type=float, default=60.0)
    ap.add_argument("--cpu-baseline-steps", type=int, default=24)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-ensemble", action="store_true",
                    help="skip the ragged column ensemble (512 columns, z_elem in {30, 45, 63, 90, 128}, 360 steps)")
    ap.add_argument("--no-adjoint", action="store_true",
                    help="skip the differentiation workload (h8 z30, 96 steps, D0 gradient of the column-integrated "
                         "temperature loss)")
    ap.add_argument("--adjoint-only", action="store_true",
                    help="run only the differentiation workload and print its figures (profiling runs)")
    ap.add_argument("--ensemble-only", action="store_true",
                    help="run only the ragged column ensemble and print its figures (profiling runs)")
    ap.add_argument("--no-variants", action="store_true",
                    help="skip the surface-flux and implicit-diffusion step timings")
    ap.add_argument("--variants-only", action="store_true",
                    help="only the surface-flux and implicit-diffusion step timings (one JSON line)")
    ap.add_argument("--variant-steps", type=int, default=30)
    ap.add_argument("--graphs-only", action="store_true",
                    help="only the eager and hipGraph-replayed step at h8 z30, h16 z63 and h30 z63 (one JSON line)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="after the measurement, run a plain 256 MiB device-to-device copy (three launches) whose "
                         "FETCH_SIZE / WRITE_SIZE counters calibrate the traffic of the step kernels")
    return ap.parse_args()


def kernel_bytes(S, nC, nF):
    """Algorithmic (compulsory) bytes per launch of each kernel of the step,
    keyed by the profile names of the library.  fi = one implicit tendency
    (rho, rho e_tot and u_3 planes).  Newton stage i (i = 1, 2, 3) reads u,
    i explicit and i - 1 implicit tendencies and writes U and (i < 3) F_I;
    the three stages are averaged, as their launches are.  The last explicit launch
    (k_explicit_final) also forms the final combination U_4 + dt sum_j c_j
    F_E_j: it reads the three earlier explicit tendencies besides its usual
    inputs and writes u_{n+1} in place of a tendency."""
    fi = 2 * nC + nF
    newton = sum((S + i * S + (i - 1) * fi) + (S + (fi if i < 3 else 0)) for i in (1, 2, 3)) / 3.0
    return {
        "k_laplacians": 4 * (S + 4 * nC),
        "k_dss(laplacians)": 4 * (2 * 4 * nC),
        "k_explicit": 4 * (S + 4 * nC + S),
        "k_explicit_final": 4 * (S + 4 * nC + 3 * S + S),
        "k_dss(tendency)": 4 * (2 * S),
        "k_newton_stage": 4 * newton,
        "k_dss(state)": 4 * (2 * S),
    }


def read_traffic(kernel):
    path = os.path.join(ROOT, "profiles", f"{kernel}_traffic.txt")
    if not os.path.exists(path):
        return None
    vals = {}
    with open(path) as f:
        for line in f:
            if ":" in line and not line.startswith("#"):
                k, v = line.split(":", 1)
                vals[k.strip()] = v.strip()
    try:
        return float(vals["traffic_bytes_per_launch"])
    except (KeyError, ValueError):
        return None


def host_threads():
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        return int(env)
    return len(os.sched_getaffinity(0))


def cpu_baseline(grid, params, numerics, Y0, dt, nsteps):
    """The CPU restatement (oracle/, Float32 build, OpenMP over columns,
    elements and node groups) stepping the same workload on the host."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from oracle import Oracle  # the CPU restatement: timing baseline only

    import numpy as np

    o = Oracle(grid, params, numerics, f32=True)
    Y = np.asarray(Y0, dtype=np.float32).copy()
    o.dss(Y)
    o.step(Y, dt, 1)  # first touch of the scratch arrays, untimed
    t0 = time.perf_counter()
    o.step(Y, dt, nsteps)
    el = time.perf_counter() - t0
    per_step = el / nsteps
    threads = host_threads()
    return {
        "value": round((dt / per_step) / 365.0, 5),
        "unit": "SYPD",
        "cores": threads,
        "kind": "port",
        "sample": f"{nsteps} full ARS343 steps of the same h{grid.h_elem} z{grid.nlev} workload (after one untimed "
                  f"step), Float32, {threads} OpenMP threads: {per_step:.3f} s/step",
    }


def cpu_baseline_c3(params, members, states, dt, nsteps):
    """The CPU restatement (oracle/, Float64, one column per call) stepping
    every member of the ensemble workload for nsteps, members spread over a
    pool of host threads (each oracle context is independent; the calls
    release the interpreter lock)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from concurrent.futures import ThreadPoolExecutor

    import numpy as np
    from oracle import Oracle  # the CPU restatement: timing baseline only

    from cadycore import geometry as G
    from cadycore.params import Numerics

    num = Numerics(hyperdiff=False, dt=dt, rayleigh=True, vert_diff=True)
    work = []
    for mem, Y0 in zip(members, states):
        g = G.columns(1, mem.z_face, params.f_plane)
        o = Oracle(g, params, num, alpha_w_col=np.array([mem.alpha_w]), D0_col=np.array([mem.D0]))
        Y = np.asarray(Y0, dtype=np.float64).copy()
        o.dss(Y)
        o.set_precomputed(Y)
        work.append((o, Y))
    threads = host_threads()
    order = sorted(range(len(work)), key=lambda m: -work[m][1].size)  # longest first
    t0 = time.perf_counter()
    with ThreadPoolExecutor(threads) as ex:
        list(ex.map(lambda m: work[m][0].step(work[m][1], dt, nsteps), order))
    el = time.perf_counter() - t0
    return {
        "value": round(nsteps / el, 2),
        "unit": "ensemble steps/s",
        "column_steps_per_s": round(nsteps * len(work) / el, 1),
        "cores": threads,
        "kind": "port",
        "sample": f"{len(work)} members x {nsteps} steps (all of the workload at 360), Float64, one column per call, "
                  f"{threads} host threads: {el:.2f} s",
    }


def ensemble_c3(torch, params, cpu=True):
    """512 independent columns of uneven vertical sizes and per-member
    parameters, 360 ARS343 steps of 10 s in one persistent launch (after the
    autotuner picked the schedule).  The state is cache-resident (0.7 MB), so
    the figure is steps per second, not an HBM fraction."""
    from cadycore.ensemble import ColumnEnsemble, c3_members
    from cadycore.params import Numerics

    num = Numerics(hyperdiff=False, dt=10.0, rayleigh=True, vert_diff=True)
    members, states = c3_members(params, n=512, seed=3)
    ens = ColumnEnsemble(members, params, num)
    Y0 = ens.state(states)
    tune = ens.autotune(Y0, num.dt, 20)
    Y = Y0.clone()
    ens.step(Y, num.dt, 10)
    Y = Y0.clone()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ens.step(Y, num.dt, 360)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    # the same launch captured in a hipGraph and replayed on a side stream
    # (the first call captures and replays once, untimed)
    side = torch.cuda.Stream()
    Yg = Y0.clone()
    torch.cuda.synchronize()
    ens.step_graph(Yg, num.dt, 360, side)
    side.synchronize()
    with torch.cuda.stream(side):
        Yg.copy_(Y0)
    side.synchronize()
    t0 = time.perf_counter()
    ens.step_graph(Yg, num.dt, 360, side)
    side.synchronize()
    el_g = time.perf_counter() - t0
    out = {
        "members": len(members),
        "levels": sorted({m.nlev for m in members}),
        "steps": 360,
        "ms": round(el * 1e3, 3),
        "ensemble_steps_per_s": round(360 / el, 1),
        "column_steps_per_s": round(360 * len(members) / el, 1),
        "graph_ms": round(el_g * 1e3, 3),
        "graph_ensemble_steps_per_s": round(360 / el_g, 1),
        "graph_bitwise_equal": bool(torch.equal(Yg, Y)),
        "schedule": ens.get_tuning(),
        "autotune_ms_per_20_steps": {f"lanes{k[0]}_order{k[1]}": round(v, 4) for k, v in tune.items()},
        "finite": bool(torch.isfinite(Y).all().item()),
        "state_MB": round(ens.S * 4 / 1e6, 3),
    }
    ens.close()
    if cpu:
        out["cpu_baseline"] = cpu_baseline_c3(params, members, states, num.dt, 360)
    return out


def adjoint_c4(torch, params):
    """The differentiation workload (global h8 z30, 96 steps of 150 s,
    checkpoints every 8): forward step, ca_step_record and the pullback of
    the squared column-integrated temperature loss (D0 gradient), best of 3.
    The adjoint/forward ratio is pullback / forward; the record costs a
    forward."""
    from cadycore.lib import DryDycore
    from cadycore.workloads import c4_case, c4_loss

    case = c4_case(params)
    g, n = case["grid"], case["nsteps"]
    d = DryDycore(g, params, case["numerics"], D0_col=case["D0"])
    Y0 = d.state(case["Y0"])
    Y = Y0.clone()

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        best = 1e9
        for _ in range(3):
            t0 = time.perf_counter()
            fn()
            torch.cuda.synchronize()
            best = min(best, time.perf_counter() - t0)
        return best

    tf = timed(lambda: d.step(Y.copy_(Y0), case["dt"], n))
    tr = timed(lambda: d.step_record(Y.copy_(Y0), case["dt"], n, case["every"]))
    Yd = Y.double().requires_grad_(True)
    L = c4_loss(Yd, case, xp=torch)
    L.backward()
    YbN = Yd.grad.float().contiguous()
    D0b = torch.zeros(g.ncol, dtype=torch.float32, device=Y.device)
    tp = timed(lambda: d.step_pullback(YbN, D0b))
    # every = 0: the forward keeps every step's stage records (96 x 37 MB at
    # this size) and the pullback only reverses; same gradient, bitwise
    tra = timed(lambda: d.step_record(Y.copy_(Y0), case["dt"], n, 0))
    D0a = torch.zeros_like(D0b)
    tpa = timed(lambda: d.step_pullback(YbN, D0a))
    out = {
        "workload": "dry global h_elem=8 Nq=4 z_elem=30, dt=150 s, 96 steps, checkpoints every 8, "
                    "D0 gradient of mean (integral T dz - target)^2",
        "forward_ms_per_step": round(tf / n * 1e3, 4),
        "record_ms_per_step": round(tr / n * 1e3, 4),
        "pullback_ms_per_step": round(tp / n * 1e3, 4),
        "pullback_forward_ratio": round(tp / tf, 3),
        "record_plus_pullback_forward_ratio": round((tr + tp) / tf, 3),
        "records_kept": {
            "record_ms_per_step": round(tra / n * 1e3, 4),
            "pullback_ms_per_step": round(tpa / n * 1e3, 4),
            "pullback_forward_ratio": round(tpa / tf, 3),
            "record_plus_pullback_forward_ratio": round((tra + tpa) / tf, 3),
            "gradient_bitwise_equal": bool(torch.equal(D0a, D0b)),
        },
        "loss": float(L.detach()),
        "finite": bool(torch.isfinite(D0b).all().item()),
    }
    d.close()
    return out


def copy_calibration(torch):
    """Three launches of a plain device copy of 256 MiB (read 256 MiB, write
    256 MiB): the reference reading for the FETCH_SIZE / WRITE_SIZE counters."""
    n = 64 * 1024 * 1024
    src = torch.ones(n, dtype=torch.float32, device="cuda:0")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()


def step_graph_sizes(torch, par, steps=50):
    """ms per step of ca_step against its hipGraph replay (ca_step_graph) at
    the sizes of the workloads: h8 z30 (the differentiation workload's grid),
    h16 z63 (the 50-step sphere case) and h30 z63.  One call of `steps`
    steps each, after a warm-up call (which also captures the graph)."""
    from cadycore import geometry as G
    from cadycore import initial as I
    from cadycore import params as P
    from cadycore.lib import DryDycore

    out = {}
    side = torch.cuda.Stream()
    for he, ze, dzb, dt in ((8, 30, 500.0, 150.0), (16, 63, 30.0, 100.0), (30, 63, 30.0, 60.0)):
        g = G.cubed_sphere(he, G.tanh_stretched_faces(ze, 30000.0, dzb), par.planet_radius, par.Omega)
        d = DryDycore(g, par, P.Numerics(dt=dt))
        Y0 = d.state(I.sphere_state(g, par, seed=2))
        Y = Y0.clone()
        res = {}
        for name in ("eager", "graph"):
            def run(n):
                if name == "eager":
                    d.step(Y, dt, n)
                    torch.cuda.synchronize()
                else:
                    d.step_graph(Y, dt, n, side)
                    side.synchronize()
            Y.copy_(Y0)
            torch.cuda.synchronize()
            run(3)
            Y.copy_(Y0)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            run(steps)
            res[name] = (time.perf_counter() - t0) / steps * 1e3
        out[f"h{he}_z{ze}"] = {"eager_ms_per_step": round(res["eager"], 4), "graph_ms_per_step": round(res["graph"], 4),
                              "speedup": round(res["eager"] / res["graph"], 3)}
        d.close()
    return out


def step_variants(torch, grid, par, Y0, dt, steps=30, warmup=5):
    """ms per step of the same workload with the optional terms on: the surface
    fluxes (ExchangeCoefficients Cd = Ch = 0.0044, the analytic surface
    temperature; on in the reference's default AtmosModel) and vertical
    diffusion (D0 = 5 m^2/s in every column) treated implicitly with one
    iteration of the approximate arrowhead solve."""
    import numpy as np

    from cadycore import initial as I
    from cadycore import params as P
    from cadycore.lib import DryDycore

    cases = {
        "surface_flux": (dict(sfc_flux=True), dict(T_sfc_col=I.surface_temperature(grid))),
        "implicit_vertical_diffusion": (dict(vert_diff=True, implicit_diffusion=True),
                                        dict(D0_col=np.full(grid.ncol, 5.0))),
        "surface_flux+implicit_vertical_diffusion": (
            dict(sfc_flux=True, vert_diff=True, implicit_diffusion=True),
            dict(T_sfc_col=I.surface_temperature(grid), D0_col=np.full(grid.ncol, 5.0))),
    }
    out = {}
    for name, (nk, hk) in cases.items():
        d = DryDycore(grid, par, P.Numerics(dt=dt, **nk), **hk)
        Y = d.state(Y0)
        d.step(Y, dt, warmup)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        d.step(Y, dt, steps)
        torch.cuda.synchronize()
        el = (time.perf_counter() - t0) / steps
        out[name] = {"ms_per_step": round(el * 1e3, 4), "SYPD": round(dt / el / 365.0, 3), "steps": steps,
                     "finite": bool(d.check_finite(Y))}
        d.close()
        del Y
    return out


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {a.gpus}); multi-GPU sharding is not built",
              file=sys.stderr)
        return 2
    import numpy as np
    import torch

    from cadycore import geometry as G
    from cadycore import initial as I
    from cadycore import params as P
    from cadycore.lib import DryDycore

    par = P.Params.from_toml()
    if a.variants_only:
        zf = G.tanh_stretched_faces(a.z_elem, 30000.0, 30.0)
        grid = G.cubed_sphere(a.h_elem, zf, par.planet_radius, par.Omega)
        out = step_variants(torch, grid, par, I.sphere_state(grid, par, seed=2), a.dt, steps=a.variant_steps)
        print(json.dumps({"workload": f"h_elem={a.h_elem} z_elem={a.z_elem} dt={a.dt:g}s", "step_variants": out}))
        return 0
    if a.graphs_only:
        print(json.dumps({"step_graph": step_graph_sizes(torch, par)}))
        return 0
    if a.adjoint_only or a.ensemble_only:
        out = adjoint_c4(torch, par) if a.adjoint_only else ensemble_c3(torch, par, cpu=False)
        print(json.dumps(out))
        return 0
    zf = G.tanh_stretched_faces(a.z_elem, 30000.0, 30.0)
    grid = G.cubed_sphere(a.h_elem, zf, par.planet_radius, par.Omega)
    num = P.Numerics(dt=a.dt)
    Y0 = I.sphere_state(grid, par, seed=2)
    dyc = DryDycore(grid, par, num)
    S = dyc.S
    nC, nF = grid.nlev * grid.ncol, (grid.nlev + 1) * grid.ncol
    Y = dyc.state(Y0)
    stream = torch.cuda.current_stream()
    if a.warmup:
        dyc.step(Y, a.dt, a.warmup)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    dyc.step(Y, a.dt, a.steps)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    ok = dyc.check_finite(Y)
    ms = el / a.steps * 1e3
    sypd = (a.dt / (el / a.steps)) / 365.0
    b_alg = sum(LAUNCHES_PER_STEP[k] * v for k, v in kernel_bytes(S, nC, nF).items())
    step_gbps = b_alg / (el / a.steps) / 1e9

    # per-kernel device time with HIP events on the launch stream (separate, untimed pass)
    nprof = min(20, max(2, a.steps // 10))
    dyc.profile_enable(True)
    dyc.step(Y, a.dt, nprof)
    prof = dyc.profile()
    dyc.profile_enable(False)
    dom = max(prof.items(), key=lambda kv: kv[1][1])
    dom_name, (dom_n, dom_ms) = dom
    avg_s = dom_ms / dom_n / 1e3
    kb = kernel_bytes(S, nC, nF)
    per_launch = kb.get(dom_name)
    achieved = per_launch / avg_s / 1e9 if per_launch else None
    traffic = read_traffic(dom_name)
    roof = {
        "kernel": dom_name,
        "bound": "hbm",
        "achieved": round(achieved, 1) if achieved else None,
        "peak": HBM_PEAK_GBPS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBPS, 4) if achieved else None,
        "traffic": traffic,
        "avg_launch_us": round(avg_s * 1e6, 2),
        "alg_bytes_per_launch": per_launch,
    }
    kernels = {}
    for k, (n, t) in prof.items():
        if not n:
            continue
        kernels[k] = {"launches_per_step": n / nprof, "us_per_launch": round(t / n * 1e3, 2),
                      "ms_per_step": round(t / nprof, 4)}
        if k in kb:
            kernels[k]["alg_GBps"] = round(kb[k] / (t / n / 1e3) / 1e9, 1)

    if a.copy_calibration:
        copy_calibration(torch)
    variants = None if a.no_variants else step_variants(torch, grid, par, Y0, a.dt, steps=a.variant_steps)
    graphs = None if a.no_variants else step_graph_sizes(torch, par)
    ens = None if a.no_ensemble else ensemble_c3(torch, par, cpu=not a.no_cpu_baseline)
    adj = None if a.no_adjoint else adjoint_c4(torch, par)

    cpu = None
    if not a.no_cpu_baseline and a.cpu_baseline_steps > 0:
        cpu = cpu_baseline(grid, par, num, Y0, a.dt, a.cpu_baseline_steps)

    line = {
        "metric": METRIC,
        "value": round(sypd, 3),
        "unit": "SYPD",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded balanced profile + jet + noise)",
        "config": {
            "workload": f"dry global cubed sphere h_elem={a.h_elem} Nq=4 z_elem={a.z_elem} (30 km, dz_bottom 30 m), "
                        f"FP32, ARS343 dt={a.dt:g}s",
            "columns": grid.ncol,
            "state_floats": S,
        },
        "achieved_step_GBps": round(step_gbps, 1),
        "step_roofline_frac": round(step_gbps / HBM_PEAK_GBPS, 4),
        "steps_per_s": round(1e3 / ms, 2),
        "adjoint_forward_ratio": adj["pullback_forward_ratio"] if adj else None,
        "gradient_forward_ratio": adj["record_plus_pullback_forward_ratio"] if adj else None,
        "finite": bool(ok),
        "roofline": roof,
        "cpu_baseline": cpu,
        "step_variants": variants,
        "step_graph": graphs,
        "c3_ensemble": ens,
        "c4_adjoint": adj,
        "kernels": kernels,
    }
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
