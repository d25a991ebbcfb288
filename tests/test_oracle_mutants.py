"""Mutation run over the oracle's explicit tendency: each mutant is a copy of
oracle/ca_oracle.c with one plausible mistake planted (a flipped sign, a
dropped term, a misplaced factor), built into a temporary directory.  The
pinning checks of test_oracle_tendency_pinning.py must fail on every mutant:
a restatement with that mistake would not pass them.

The mistakes are planted on the CPU in the oracle only; nothing here touches
the native library.
"""
import os
import shutil
import subprocess
import tempfile

import pytest

import test_oracle_tendency_pinning as T
from oracle import Oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "oracle", "ca_oracle.c")

# (name, text in ca_oracle.c, replacement, the check expected to catch it):
# each text occurs exactly once
MUTANTS = [
    ('split_div(rho u, 1) sign',
     'RHO(Yt)[o + col] -= wd[col] / 2 + (wd[col] + 0) / 2;',
     'RHO(Yt)[o + col] += wd[col] / 2 + (wd[col] + 0) / 2;',
     'mms_horizontal'),
    ('split_div(rho u, 1) half dropped',
     'RHO(Yt)[o + col] -= wd[col] / 2 + (wd[col] + 0) / 2;',
     'RHO(Yt)[o + col] -= wd[col] / 2;',
     'mms_horizontal'),
    ('split_div(rho u, h): rho u . grad h dropped',
     '(c->h[o + col] * wd[col] + (m1[col] * gh1[col] + m2[col] * gh2[col])) / 2;',
     '(c->h[o + col] * wd[col]) / 2;',
     'mms_horizontal'),
    ('split_div(rho u, h): h wdiv(rho u) sign',
     '(c->h[o + col] * wd[col] + (m1[col] * gh1[col] + m2[col] * gh2[col])) / 2;',
     '(-c->h[o + col] * wd[col] + (m1[col] * gh1[col] + m2[col] * gh2[col])) / 2;',
     'mms_horizontal'),
    ("Exner PGF: Pi grad theta' sign",
     'UH1(Yt)[o + col] -= g1[col] + cp * (th[col] * p1[col] + q1[col] - Pi[col] * r1[col]) / 2;',
     'UH1(Yt)[o + col] -= g1[col] + cp * (th[col] * p1[col] + q1[col] + Pi[col] * r1[col]) / 2;',
     'mms_horizontal'),
    ("Exner PGF: grad(theta' Pi) dropped",
     'UH2(Yt)[o + col] -= g2[col] + cp * (th[col] * p2[col] + q2[col] - Pi[col] * r2[col]) / 2;',
     'UH2(Yt)[o + col] -= g2[col] + cp * (th[col] * p2[col] - Pi[col] * r2[col]) / 2;',
     'mms_horizontal'),
    ('grad(K + Phi - Phi_r) sign',
     'UH1(Yt)[o + col] -= g1[col] + cp *',
     'UH1(Yt)[o + col] -= -g1[col] + cp *',
     'mms_horizontal'),
    ('reference state: Phi_r dropped from the gradient',
     'A[col] = c->K[o + col] + c->Phic[k] - Phi_ref(c, P);',
     'A[col] = c->K[o + col] + c->Phic[k];',
     'mms_horizontal'),
    ('vertical Coriolis f^3 sign',
     'real a3 = (real)c->g.f3_num[col] / c->Jvc[k] + om3[q];',
     'real a3 = -(real)c->g.f3_num[col] / c->Jvc[k] + om3[q];',
     'mms_horizontal'),
    ('vorticity omega^3 sign',
     'real a3 = (real)c->g.f3_num[col] / c->Jvc[k] + om3[q];',
     'real a3 = (real)c->g.f3_num[col] / c->Jvc[k] - om3[q];',
     'mms_horizontal'),
    ('horizontal Coriolis f^12 sign',
     'real A1 = fa + w12a[q], A2 = fb + w12b[q];',
     'real A1 = -fa + w12a[q], A2 = -fb + w12b[q];',
     'mms_vertical'),
    ('curl_v(u_h) sign',
     'a = -(u2[IDX(f, col)] - u2[IDX(f - 1, col)]) / Jf;',
     'a = (u2[IDX(f, col)] - u2[IDX(f - 1, col)]) / Jf;',
     'mms_vertical'),
    ('wcurl(u_3) in omega^12 sign',
     'w12a[o + col] += x[col];',
     'w12a[o + col] -= x[col];',
     'mms_with_w'),
    ('grad_v(K) on u_3 sign',
     'U3(Yt)[q] -= cr3 + gK;',
     'U3(Yt)[q] -= cr3 - gK;',
     'mms_vertical'),
    ('(f12 + omega12) x I(CT12 u) on u_3 dropped',
     'U3(Yt)[q] -= cr3 + gK;',
     'U3(Yt)[q] -= gK;',
     'mms_vertical'),
    ('(f12 + omega12) x u^3 on u_h sign',
     'UH1(Yt)[q] -= ia / rJ + x1;',
     'UH1(Yt)[q] -= -ia / rJ + x1;',
     'mms_with_w'),
    ('vector Laplacian: curl-curl sign',
     'c->L1[o + col] = gd1[col] - wc1[col];',
     'c->L1[o + col] = gd1[col] + wc1[col];',
     'hyperdiffusion'),
    ('divergence damping on the curl part',
     'real d1 = dd * gd1[col] - wc1[col];',
     'real d1 = gd1[col] - dd * wc1[col];',
     'hyperdiffusion'),
    ('u_h hyperdiffusion sign',
     'UH2(Yt)[o + col] -= nv * d2;',
     'UH2(Yt)[o + col] += nv * d2;',
     'hyperdiffusion'),
    ('scalar hyperdiffusion: rho outside the divergence dropped',
     'v1[col] = rho[o + col] * a1;\n                    v2[col] = rho[o + col] * a2;',
     'v1[col] = a1;\n                    v2[col] = a2;',
     'hyperdiffusion'),
    ('s_d reference state dropped from the scalar Laplacian',
     'sd[col] = (cpd * (c->T[o + col] - T0) + c->Phic[k]) - sdr;',
     'sd[col] = (cpd * (c->T[o + col] - T0) + c->Phic[k]);',
     'hyperdiffusion'),
    ('u_3 hyperdiffusion sign',
     'U3(Yt)[IDX(f, col)] -= nv * v;',
     'U3(Yt)[IDX(f, col)] += nv * v;',
     'hyperdiffusion'),
    ('vertical diffusion of rho e_tot sign',
     'RHOE(Yt)[q] -= (F[IDX(k + 1, col)] - F[IDX(k, col)]) / Jc;',
     'RHOE(Yt)[q] += (F[IDX(k + 1, col)] - F[IDX(k, col)]) / Jc;',
     'vdiff_column'),
    ('vertical diffusion: face rho from the lower centre',
     'return ((rho[IDX(f - 1, col)] + rho[IDX(f, col)]) / 2) / ((1 / Ka + 1 / Kb) / 2);',
     'return rho[IDX(f - 1, col)] / ((1 / Ka + 1 / Kb) / 2);',
     'vdiff_sphere'),
    ('vertical diffusion: arithmetic face mean of K',
     '((1 / Ka + 1 / Kb) / 2)',
     '(2 / (Ka + Kb))',
     'vdiff_closures'),
    ('DecayWithHeightDiffusion: decay height ignored',
     '(real)exp(-(double)(c->zc[k] - c->zf[0]) / c->n.vdiff_H)',
     '(real)1',
     'vdiff_closures'),
    ('VerticalDiffusion: taper sign',
     'return KE * (real)exp(-(double)(x * x));',
     'return KE * (real)exp((double)(x * x));',
     'vdiff_closures'),
    ('VerticalDiffusion: z_a the whole lowest layer',
     'real za = (c->zf[1] - c->zf[0]) / 2;',
     'real za = (c->zf[1] - c->zf[0]);',
     'vdiff_closures'),
    ('vertical diffusion of u_h sign',
     'Fu1[q] = fac * (ub[0] - ua[0]) * inv;',
     'Fu1[q] = fac * (ua[0] - ub[0]) * inv;',
     'vdiff_column'),
    ('vertical diffusion: deep-atmosphere face Jacobian dropped',
     'F[q] = c->Jf[q] * (g33 * (-(rK * (sdb - sda))));',
     'F[q] = c->Jvf[f] * (g33 * (-(rK * (sdb - sda))));',
     'vdiff_sphere'),
    ('Rayleigh u_h sign',
     'UH1(Yt)[IDX(k, col)] += -beta * u1[IDX(k, col)];',
     'UH1(Yt)[IDX(k, col)] += beta * u1[IDX(k, col)];',
     'rayleigh_uh'),
    ('surface heat flux sign',
     'const real shf = -rho_sfc * Ch * speed * (dse_int - dse_sfc);',
     'const real shf = rho_sfc * Ch * speed * (dse_int - dse_sfc);',
     'sfc_column'),
    ('surface density taken as the interior density',
     'const real rho_sfc = p_sfc / (R * T_sfc);',
     'const real rho_sfc = rho_int;',
     'sfc_column'),
    ('surface density extrapolated at the interior temperature',
     'const real p_sfc = p_int * exp(grav * dz / (R * (T_int + T_sfc) / 2));',
     'const real p_sfc = p_int * exp(grav * dz / (R * T_int));',
     'sfc_sphere'),
    ('dry static energy: geopotential difference dropped',
     'const real dse_int = cpd * (T_int - T0) + Phi_int, dse_sfc',
     'const real dse_int = cpd * (T_int - T0) + Phi_sfc, dse_sfc',
     'sfc_column'),
    ('gustiness floor dropped',
     'if (speed < gust) speed = gust;',
     '',
     'sfc_column'),
    ('surface drag swaps Cd and Ch',
     'const real tau_x = -rho_sfc * Cd * speed * u_int, tau_y = -rho_sfc * Cd * speed * v_int;',
     'const real tau_x = -rho_sfc * Ch * speed * u_int, tau_y = -rho_sfc * Ch * speed * v_int;',
     'sfc_column'),
    ('surface momentum flux left contravariant',
     'const real f1 = g11 * fc1 + g12 * fc2, f2 = g12 * fc1 + g22 * fc2; /* C12 */',
     'const real f1 = fc1, f2 = fc2;',
     'sfc_sphere'),
    ('implicit diffusion: cp_d / cv_d dropped from the (rho e, rho e) block',
     'c->Aee_d[q] = dtg * dd * (cp / (cv * rho[q])) - 1;',
     'c->Aee_d[q] = dtg * dd * (1 / rho[q]) - 1;',
     'idiff_blocks'),
    ('implicit diffusion: (u_h, u_h) block sign',
     'c->Ahh_lo[q] = dtg * (1 / rho[q]) * dlo;',
     'c->Ahh_lo[q] = -dtg * (1 / rho[q]) * dlo;',
     'idiff_blocks'),
    ('implicit diffusion: tendency left in the explicit tendency as well',
     'if (c->n.vert_diff && !c->n.implicit_diffusion) vertical_diffusion_tendency(c, Yt, Y);',
     'if (c->n.vert_diff) vertical_diffusion_tendency(c, Yt, Y);',
     'idiff_move'),
    ('approximate solve: Schur complement inv(A_ee) term sign',
     'sx -= CB(Aue0, f) * m[a] + CB(Aue1, f) * m[b];',
     'sx += CB(Aue0, f) * m[a] + CB(Aue1, f) * m[b];',
     'idiff_solve'),
    ('approximate solve: A21 inv(A11) b1 sign',
     'bw[f] -= CB(Aur0, f) * (-br[f - 1]) + CB(Aur1, f) * (-br[f]) + CB(Aue0, f) * ye[f - 1] + CB(Aue1, f) * ye[f];',
     'bw[f] += CB(Aur0, f) * (-br[f - 1]) + CB(Aur1, f) * (-br[f]) + CB(Aue0, f) * ye[f - 1] + CB(Aue1, f) * ye[f];',
     'idiff_solve'),
    ('approximate solve: back substitution without inv(A_ee)',
     'dY[3 * nC + q] = m[k];',
     'dY[3 * nC + q] = -g[k];',
     'idiff_solve'),
    ('surface drag divided by the surface density',
     'UH1(Yt)[q] -= div_1 / rho_int;',
     'UH1(Yt)[q] -= div_1 / rho_sfc;',
     'sfc_column'),
]


@pytest.fixture(scope="module")
def workdir():
    d = tempfile.mkdtemp(prefix="ora_mut_")
    yield d
    shutil.rmtree(d, ignore_errors=True)


def _build(workdir, i, text, repl):
    src = open(SRC).read()
    assert src.count(text) == 1, f"mutant {i}: the planted text must occur once"
    path_c = os.path.join(workdir, f"mut{i}.c")
    with open(path_c, "w") as f:
        f.write(src.replace(text, repl))
    shutil.copy(os.path.join(ROOT, "oracle", "ca_oracle.h"), os.path.join(workdir, "ca_oracle.h"))
    so = os.path.join(workdir, f"libmut{i}.so")
    subprocess.check_call(["gcc", "-O1", "-std=c11", "-fPIC", "-shared", "-w", "-o", so, path_c, "-lm"])
    return so


@pytest.mark.parametrize("i", range(len(MUTANTS)), ids=[m[0] for m in MUTANTS])
def test_pinning_checks_catch_mutant(workdir, i):
    name, text, repl, first = MUTANTS[i]
    so = _build(workdir, i, text, repl)

    def mk(grid, params, numerics, **kw):
        return Oracle(grid, params, numerics, lib_path=so, **kw)

    caught = []
    order = [first] + [c for c in T.ALL_CHECKS if c != first]
    for cname in order:
        check = T.ALL_CHECKS[cname]
        try:
            check(mk)
        except AssertionError:
            caught.append(cname)
            break
    assert caught, f"no pinning check fails on the mutant '{name}'"
