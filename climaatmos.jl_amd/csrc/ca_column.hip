// ca_column.hip -- column kernels of the dry-dycore step: precomputed
// quantities, implicit tendency, Jacobian bands, the Schur-complement solve,
// the post-Newton upwind correction, and the fused per-column Newton stage.
//
// Restated from (reference paths):
//   src/cache/precomputed_quantities.jl:698-831        (precomputed, filter)
//   src/prognostic_equations/implicit/implicit_tendency.jl:185-298, 322-339
//   src/prognostic_equations/implicit/manual_sparse_jacobian.jl:713-870, 579-584
//   src/utils/abbreviations.jl:92-256                  (vertical operators + BCs)
#include "ca_kernels.h"

// ---------------------------------------------------------------- helpers
// Filtered view of u_3: the impenetrability filter of
// set_velocity_at_surface!/set_velocity_at_top! (precomputed_quantities.jl:
// 486-554) sets u_3 = 0 on the bottom and top faces (no topography).
__device__ __forceinline__ float u3_at(const DevGeom &g, const float *u3, int f, int col) {
    return (f <= 0 || f >= g.nlev) ? 0.0f : u3[(int64_t)f * g.ncol + col];
}

// Centre quantities of a column at level k (compute_kinetic, utilities.jl:194-212,
// and the dry thermodynamics of precomputed_quantities.jl:786-815).
__device__ __forceinline__ CV cv_make(const DevGeom &g, const DevNum &n, int k, int col, float rho, float u1,
                                      float u2, float rhoe, float u3a, float u3b) {
    CV c;
    c.rho = rho;
    c.u1 = u1;
    c.u2 = u2;
    c.rhoe = rhoe;
    float s = g.sc[k];
    float is2 = g.is2c[k];
    float ct1, ct2;
    ca_contra(g, col, is2, u1, u2, ct1, ct2);
    c.ct1 = ct1;
    c.ct2 = ct2;
    float ga = g.g33f[k];
    float gb = g.g33f[k + 1];
    float vv = 0.5f * (u3a * ga * u3a + u3b * gb * u3b);
    c.K = 0.5f * ((u1 * ct1 + u2 * ct2) + vv);
    // thermodynamics in float64: the implicit pressure gradient is a small
    // difference of large hydrostatic terms, so T and p are kept to double
    // precision for it (the state itself stays float32)
    double e_tot = (double)rhoe * ca_drcp((double)rho);
    double e_int = e_tot - (double)c.K - g.Phic_d[k];
    double Td = n.dT_0 + (e_int + n.dR_d * n.dT_0) * n.dicv_d;
    Td = Td > n.dTmin ? Td : n.dTmin;
    double pd = (double)rho * n.dR_d * Td;
    c.Td = Td;
    c.pd = pd;
    c.Pid = ca_dexner(n, pd * n.dp0inv);
    c.T = (float)Td;
    c.h = (float)(e_tot + n.dR_d * Td);
    c.p = (float)pd;
    c.Pi = (float)c.Pid;
    c.J = g.Jh[col] * s * s * g.Jvc[k];
    c.iJ = g.iJh[col] * g.isJc[k];
    return c;
}

// h_tot = e_tot + R_d T of a centre with cv_make's arithmetic (bitwise the
// same h), without the pressure and Exner function it does not need
__device__ __forceinline__ float cv_enthalpy(const DevGeom &g, const DevNum &n, int k, int col, float rho, float u1,
                                             float u2, float rhoe, float u3a, float u3b) {
    float ct1, ct2;
    ca_contra(g, col, g.is2c[k], u1, u2, ct1, ct2);
    const float vv = 0.5f * (u3a * g.g33f[k] * u3a + u3b * g.g33f[k + 1] * u3b);
    const float K = 0.5f * ((u1 * ct1 + u2 * ct2) + vv);
    const double e_tot = (double)rhoe * ca_drcp((double)rho);
    const double e_int = e_tot - (double)K - g.Phic_d[k];
    double Td = n.dT_0 + (e_int + n.dR_d * n.dT_0) * n.dicv_d;
    Td = Td > n.dTmin ? Td : n.dTmin;
    return (float)(e_tot + n.dR_d * Td);
}

__device__ __forceinline__ CV cv_load(const DevGeom &g, const DevNum &n, const float *Y, int k, int col) {
    const int64_t q = (int64_t)k * g.ncol + col;
    const float *u3 = Y + 4 * g.nC;
    return cv_make(g, n, k, col, Y[q], Y[g.nC + q], Y[2 * g.nC + q], Y[3 * g.nC + q], u3_at(g, u3, k, col),
                   u3_at(g, u3, k + 1, col));
}

// face value of the upwinded product U(u3, chi) at an interior face f
// (abbreviations.jl:227-256): 0 central, 1 first-order upwind, 3 Lin (1994)
// van Leer with monotone local-extrema constraint and first-order one-sided
// stencils on faces 1 and nlev - 1.  c2..c5 = chi at centres f-2 .. f+1.
__device__ __forceinline__ float upwind_product(int scheme, int f, int nlev, float v, float dt, float c2, float c3,
                                                float c4, float c5) {
    if (scheme == 0) return v * (0.5f * (c3 + c4));
    if (scheme == 1 || f == 1 || f == nlev - 1) return v >= 0.0f ? v * c3 : v * c4;
    float cour = v * dt;
    if (v >= 0.0f) return v * (c3 + (1.0f - cour) * ca_vl_slope(c2, c3, c4) * 0.5f);
    return v * (c4 - (1.0f + cour) * ca_vl_slope(c3, c4, c5) * 0.5f);
}

// Implicit u_3 tendency at an interior face from the centres below (a) and
// above (b): -(grad Phi - grad Phi_r + cp_d I^f(theta_v - theta_vr) grad Pi)
// (implicit_tendency.jl:287-293) in float64 from the centres' float64 Exner
// values (relative error ~1e-13, ca_dexner): the differences Pi_b - Pi_a and
// Pi_b^7 - Pi_a^7 keep ~1e-11 of their value, far below the float32 result.
// ln(Pi_b / Pi_a) = 2 atanh(s), s = (Pi_b - Pi_a) / (Pi_b + Pi_a), by its odd
// series to s^19: |s| <= 0.13 for layers up to 6 km deep, where the
// truncation is below 1e-17 relative.
__device__ __forceinline__ double ca_dpow7(double x) {
    const double x2 = x * x, x4 = x2 * x2;
    return x4 * x2 * x;
}
__device__ __forceinline__ float pgf_u3(const DevGeom &g, const DevNum &n, int f, const CV &a, const CV &b) {
    const double Pa = a.Pid, Pb = b.Pid;
    const double dPi = Pb - Pa;
    const double Pa7 = ca_dpow7(Pa), Pb7 = ca_dpow7(Pb);
    const double s = dPi * ca_drcp(Pa + Pb), s2 = s * s;
    double t = 1.0 / 19.0;
#pragma unroll
    for (int k = 17; k >= 1; k -= 2) t = fma(t, s2, 1.0 / k);
    const double lr = 2.0 * s * t;
    const double dPhr = -n.dcp_d * (n.dTmr * lr + (n.dTsr - n.dTmr) * (1.0 / 7.0) * (Pb7 - Pa7));
    const double tha = (a.Td - (n.dTmr + (n.dTsr - n.dTmr) * Pa7)) * ca_drcp(Pa);
    const double thb = (b.Td - (n.dTmr + (n.dTsr - n.dTmr) * Pb7)) * ca_drcp(Pb);
    return (float)(-((g.dPhif_d[f] - dPhr) + n.dcp_d * (0.5 * (tha + thb)) * dPi));
}

__device__ __forceinline__ float beta_w(const DevGeom &g, const DevNum &n, int f, int col) {
    if (!n.rayleigh) return 0.0f;
    float aw = g.aw_col ? g.aw_col[col] : n.alpha_w;
    return aw * g.zeta_f[f];
}

// ------------------------------------------------------------------ kernels

// set_implicit_precomputed_quantities!: filter Y and store K, T, p, h_tot.
__global__ void k_precompute(DevGeom g, DevNum n, float *Y, float *K, float *T, float *p, float *h) {
    int col = blockIdx.x * blockDim.x + threadIdx.x;
    int k = blockIdx.y;
    if (col >= g.ncol) return;
    CV c = cv_load(g, n, Y, k, col);
    int64_t q = (int64_t)k * g.ncol + col;
    K[q] = c.K;
    T[q] = c.T;
    p[q] = c.p;
    h[q] = c.h;
    float *u3 = Y + 4 * g.nC;
    if (k == 0) u3[col] = 0.0f;
    if (k == g.nlev - 1) u3[(int64_t)g.nlev * g.ncol + col] = 0.0f;
}

// implicit_tendency! (dry subset): thread per (level l in [0, nlev], column);
// l < nlev writes the centre-l rho / rho e_tot tendency, every l the face-l
// u_3 tendency.  u_h tendencies are zero.
__global__ void k_implicit_tendency(DevGeom g, DevNum n, float *Yt, const float *Y) {
    int col = blockIdx.x * blockDim.x + threadIdx.x;
    int l = blockIdx.y;
    if (col >= g.ncol) return;
    const int nl = g.nlev;
    const float *u3 = Y + 4 * g.nC;
    if (l < nl) {
        // centre l: -advdiv(I^f(rho J)/J_f u^3 chi), zero flux on boundary faces
        CV c = cv_load(g, n, Y, l, col);
        float Fm = 0.0f, Fp = 0.0f, Hm = 0.0f, Hp = 0.0f;
        if (l >= 1) {
            CV a = cv_load(g, n, Y, l - 1, col);
            float fu3 = u3_at(g, u3, l, col) / (g.Jvf[l] * g.Jvf[l]);
            float rJ = 0.5f * (a.rho * a.J + c.rho * c.J);
            Fm = rJ * fu3;
            Hm = Fm * (0.5f * (a.h + c.h));
        }
        if (l + 1 <= nl - 1) {
            CV b = cv_load(g, n, Y, l + 1, col);
            float fu3 = u3_at(g, u3, l + 1, col) / (g.Jvf[l + 1] * g.Jvf[l + 1]);
            float rJ = 0.5f * (c.rho * c.J + b.rho * b.J);
            Fp = rJ * fu3;
            Hp = Fp * (0.5f * (c.h + b.h));
        }
        int64_t q = (int64_t)l * g.ncol + col;
        Yt[q] = -(Fp - Fm) / c.J;
        Yt[g.nC + q] = 0.0f;
        Yt[2 * g.nC + q] = 0.0f;
        Yt[3 * g.nC + q] = -(Hp - Hm) / c.J;
    }
    float t3 = 0.0f;
    if (l >= 1 && l <= nl - 1) {
        CV a = cv_load(g, n, Y, l - 1, col);
        CV b = cv_load(g, n, Y, l, col);
        t3 = pgf_u3(g, n, l, a, b) - beta_w(g, n, l, col) * u3_at(g, u3, l, col);
    }
    Yt[4 * g.nC + (int64_t)l * g.ncol + col] = t3;
}

// update_advection_jacobian! (dry, no topography): band entries scaled by dtg.
// Centre rows: A(rho,u3) and A(rhoe,u3) on faces (k, k+1).  Face rows:
// A(u3,rho), A(u3,rhoe), A(u3,uh) on centres (f-1, f); A(u3,u3) tridiagonal
// including -I and the Rayleigh term.
__device__ __forceinline__ void jac_center_row(const DevGeom &g, int k, const CV *cm, const CV &c, const CV *cp,
                                               float dtg, float &a0, float &a1, float &e0, float &e1) {
    a0 = a1 = e0 = e1 = 0.0f;
    if (cm) {
        float rJ = 0.5f * (cm->rho * cm->J + c.rho * c.J);
        float g33 = g.g33f[k];
        float adv = rJ * c.iJ;
        a0 = dtg * adv * g33;
        e0 = a0 * (0.5f * (cm->h + c.h));
    }
    if (cp) {
        float rJ = 0.5f * (c.rho * c.J + cp->rho * cp->J);
        float g33 = g.g33f[k + 1];
        float adv = -rJ * c.iJ;
        a1 = dtg * adv * g33;
        e1 = a1 * (0.5f * (c.h + cp->h));
    }
}

struct FaceRow {
    float r0, r1, e0, e1, h0a, h0b, h1a, h1b, lo, d, up;
};

__device__ __forceinline__ FaceRow jac_face_row(const DevGeom &g, const DevNum &n, int f, int col, const CV &a,
                                                const CV &b, float u3m, float u3f, float u3p, float dtg) {
    FaceRow r;
    float rf = 0.5f * (a.rho + b.rho);
    float pg_m = ca_rcp(rf), pg_p = -pg_m;
    float kap = n.kap;
    // d p / d rho at fixed rho e_tot (manual_sparse_jacobian.jl:816-818); the
    // (R_d - kappa cv_d) T term vanishes for dry air.
    float dpa = kap * (n.T_0 * n.cp_d - a.K - g.Phic[f - 1]);
    float dpb = kap * (n.T_0 * n.cp_d - b.K - g.Phic[f]);
    float thva = a.T * ca_rcp(a.Pi), thvb = b.T * ca_rcp(b.Pi);
    const float dPi = (float)(b.Pid - a.Pid);  // float64 Exner values: the difference keeps ~1e-11
    float buoy = n.cp_d * (0.5f * (thva + thvb)) * dPi * pg_m;
    r.r0 = dtg * (pg_m * dpa + 0.5f * buoy);
    r.r1 = dtg * (pg_p * dpb + 0.5f * buoy);
    r.e0 = dtg * (pg_m * kap);
    r.e1 = dtg * (pg_p * kap);
    float ka = -kap * a.rho, kb = -kap * b.rho;
    r.h0a = dtg * pg_m * ka * a.ct1;
    r.h0b = dtg * pg_m * ka * a.ct2;
    r.h1a = dtg * pg_p * kb * b.ct1;
    r.h1b = dtg * pg_p * kb * b.ct2;
    float gm = g.g33f[f - 1];
    float g0 = g.g33f[f];
    float gp = g.g33f[f + 1];
    float beta = beta_w(g, n, f, col);
    r.lo = dtg * (pg_m * ka * (0.5f * u3m * gm));
    r.d = dtg * (pg_m * ka * (0.5f * u3f * g0) + pg_p * kb * (0.5f * u3f * g0) - beta) - 1.0f;
    r.up = dtg * (pg_p * kb * (0.5f * u3p * gp));
    return r;
}

__global__ void k_update_jacobian(DevGeom g, DevNum n, const float *Y, float dtg, Bands B) {
    int col = blockIdx.x * blockDim.x + threadIdx.x;
    int l = blockIdx.y;
    if (col >= g.ncol) return;
    const int nl = g.nlev;
    const float *u3 = Y + 4 * g.nC;
    int64_t q = (int64_t)l * g.ncol + col;
    if (l < nl) {
        CV c = cv_load(g, n, Y, l, col);
        CV cm, cp;
        bool hm = l >= 1, hp = l + 1 <= nl - 1;
        if (hm) cm = cv_load(g, n, Y, l - 1, col);
        if (hp) cp = cv_load(g, n, Y, l + 1, col);
        float a0, a1, e0, e1;
        jac_center_row(g, l, hm ? &cm : nullptr, c, hp ? &cp : nullptr, dtg, a0, a1, e0, e1);
        B.Aru0[q] = a0;
        B.Aru1[q] = a1;
        B.Aeu0[q] = e0;
        B.Aeu1[q] = e1;
    }
    FaceRow r;
    if (l >= 1 && l <= nl - 1) {
        CV a = cv_load(g, n, Y, l - 1, col);
        CV b = cv_load(g, n, Y, l, col);
        r = jac_face_row(g, n, l, col, a, b, u3_at(g, u3, l - 1, col), u3_at(g, u3, l, col),
                         u3_at(g, u3, l + 1, col), dtg);
    } else {
        r.r0 = r.r1 = r.e0 = r.e1 = r.h0a = r.h0b = r.h1a = r.h1b = r.lo = r.up = 0.0f;
        r.d = -dtg * beta_w(g, n, l, col) - 1.0f;
    }
    B.Aur0[q] = r.r0;
    B.Aur1[q] = r.r1;
    B.Aue0[q] = r.e0;
    B.Aue1[q] = r.e1;
    B.Auh0a[q] = r.h0a;
    B.Auh0b[q] = r.h0b;
    B.Auh1a[q] = r.h1a;
    B.Auh1b[q] = r.h1b;
    B.lo[q] = r.lo;
    B.d[q] = r.d;
    B.up[q] = r.up;
}

// ldiv!: Schur complement on u_3 (tridiagonal) + back-substitution,
// BlockArrowheadSolve(rho, rho e_tot; alg2 = BlockLowerTriangularSolve(u_h))
// (manual_sparse_jacobian.jl:579-584).  One thread per column (Thomas).
__global__ void k_ldiv(DevGeom g, Bands B, float *dY, const float *R, float *scr) {
    int col = blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= g.ncol) return;
    const int nl = g.nlev;
    const int64_t nC = g.nC, nc = g.ncol;
    float *cpr = scr, *dpr = scr + g.nF;
    float cprev = 0.0f, dprev = 0.0f;
    for (int f = 0; f <= nl; ++f) {
        int64_t q = (int64_t)f * nc + col;
        float L = B.lo[q], D = B.d[q], U = B.up[q];
        float r = R[4 * nC + q];
        if (f >= 1 && f <= nl - 1) {
            int64_t a = q - nc, b = q;
            L += B.Aur0[q] * B.Aru0[a] + B.Aue0[q] * B.Aeu0[a];
            D += B.Aur0[q] * B.Aru1[a] + B.Aue0[q] * B.Aeu1[a] + B.Aur1[q] * B.Aru0[b] + B.Aue1[q] * B.Aeu0[b];
            U += B.Aur1[q] * B.Aru1[b] + B.Aue1[q] * B.Aeu1[b];
            r += B.Aur0[q] * R[a] + B.Aur1[q] * R[b] + B.Aue0[q] * R[3 * nC + a] + B.Aue1[q] * R[3 * nC + b];
            r += B.Auh0a[q] * R[nC + a] + B.Auh0b[q] * R[2 * nC + a] + B.Auh1a[q] * R[nC + b] +
                 B.Auh1b[q] * R[2 * nC + b];
        }
        float m = D - L * cprev;
        float cc = U / m;
        float dd = (r - L * dprev) / m;
        cpr[q] = cc;
        dpr[q] = dd;
        cprev = cc;
        dprev = dd;
    }
    float xn = 0.0f;
    for (int f = nl; f >= 0; --f) {
        int64_t q = (int64_t)f * nc + col;
        float x = dpr[q] - cpr[q] * xn;
        dY[4 * nC + q] = x;
        if (f < nl) {
            dY[q] = B.Aru0[q] * x + B.Aru1[q] * xn - R[q];
            dY[3 * nC + q] = B.Aeu0[q] * x + B.Aeu1[q] * xn - R[3 * nC + q];
            dY[nC + q] = -R[nC + q];
            dY[2 * nC + q] = -R[2 * nC + q];
        }
        xn = x;
    }
}

// correct_implicit_advection_tendency!: Yt.rhoe = vtt(upwind) - vtt(central),
// everything else zero.  Thread per (centre k, column).
__global__ void k_correct(DevGeom g, DevNum n, float *Yt, const float *Y) {
    int col = blockIdx.x * blockDim.x + threadIdx.x;
    int k = blockIdx.y;
    if (col >= g.ncol) return;
    const int nl = g.nlev;
    const float *u3 = Y + 4 * g.nC;
    float hv[6], rJv[6];
    // centres k-3 .. k+2 (as available); h relative to the lowest cell (the
    // shift cancels between the upwind and central products)
    const float href = cv_load(g, n, Y, 0, col).h;
    for (int t = 0; t < 6; ++t) {
        int kk = k - 3 + t;
        if (kk >= 0 && kk < nl) {
            CV c = cv_load(g, n, Y, kk, col);
            hv[t] = c.h - href;
            rJv[t] = c.rho * c.J;
        } else {
            hv[t] = 0.0f;
            rJv[t] = 0.0f;
        }
    }
    float dF[2];
    for (int s = 0; s < 2; ++s) {
        int f = k + s;  // faces k, k+1
        float d = 0.0f;
        if (f >= 1 && f <= nl - 1 && n.upw != 0) {
            int t = f - (k - 3);  // index of centre f in the window
            float v = u3_at(g, u3, f, col) / (g.Jvf[f] * g.Jvf[f]);
            float rJ = 0.5f * (rJv[t - 1] + rJv[t]);
            float up = upwind_product(n.upw, f, nl, v, n.dt, hv[t - 2], hv[t - 1], hv[t], t + 1 < 6 ? hv[t + 1] : 0.0f);
            float ce = upwind_product(0, f, nl, v, n.dt, 0.0f, hv[t - 1], hv[t], 0.0f);
            d = rJ * up - rJ * ce;
        }
        dF[s] = d;
    }
    int64_t q = (int64_t)k * g.ncol + col;
    float J = g.Jh[col] * g.sc[k] * g.sc[k] * g.Jvc[k];
    Yt[q] = 0.0f;
    Yt[g.nC + q] = 0.0f;
    Yt[2 * g.nC + q] = 0.0f;
    Yt[3 * g.nC + q] = -(dF[1] - dF[0]) / J;
    float *t3 = Yt + 4 * g.nC;
    t3[q] = 0.0f;
    if (k == nl - 1) t3[q + g.ncol] = 0.0f;
}

// ------------------------------------------------------ fused Newton stage
// One implicit ARS343 stage for one column: assemble the stage state
// U~ = sum_s coef_s src_s (the sources are DSSed, so U~ is continuous), apply
// the impenetrability filter, take one Newton iteration with a fresh Jacobian
// (R = dtg T_imp(U~) since U starts at U~), apply the post-Newton upwind
// correction, and write U and the implicit tendency F_I = (U - U~) / dtg
// (rho, rho e_tot, u_3 planes; the u_h component is identically zero).
// Per-face Thomas coefficients and per-centre back-substitution rows go
// through a column scratch buffer.
__global__ void __launch_bounds__(128) k_newton_stage(DevGeom g, DevNum n, StageSrc src, float dtg, float *U,
                                                      float *FI, float *scr) {
    int col = blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= g.ncol) return;
    const int nl = g.nlev;
    const int64_t nC = g.nC, nc = g.ncol;
    // ---- pass A: assemble U~ into U
    for (int k = 0; k <= nl; ++k) {
        int64_t q = (int64_t)k * nc + col;
        if (k < nl) {
            for (int comp = 0; comp < 4; ++comp) {
                float s = 0.0f;
                for (int t = 0; t < src.nfull; ++t) s = fmaf(src.cf[t], src.full[t][comp * nC + q], s);
                if (comp == 0 || comp == 3) {
                    int ci = comp == 0 ? 0 : 1;
                    for (int t = 0; t < src.nfi; ++t) s = fmaf(src.ci[t], src.fi[t][ci * nC + q], s);
                }
                U[comp * nC + q] = s;
            }
        }
        float s = 0.0f;
        if (k >= 1 && k <= nl - 1) {
            for (int t = 0; t < src.nfull; ++t) s = fmaf(src.cf[t], src.full[t][4 * nC + q], s);
            for (int t = 0; t < src.nfi; ++t) s = fmaf(src.ci[t], src.fi[t][2 * nC + q], s);
        }
        U[4 * nC + q] = s;
    }
    // ---- pass B: residual, Jacobian rows, Schur complement, Thomas forward sweep
    // scratch planes: cpr, dpr (faces); a0, a1, e0, e1, Rr, Re (centres)
    float *cpr = scr, *dpr = scr + g.nF;
    float *sa0 = scr + 2 * g.nF, *sa1 = sa0 + nC, *se0 = sa1 + nC, *se1 = se0 + nC, *sRr = se1 + nC, *sRe = sRr + nC;
    const float *u3 = U + 4 * nC;
    CV cm, c0, cp;
    c0 = cv_load(g, n, U, 0, col);
    if (nl > 1) cp = cv_load(g, n, U, 1, col);
    float pa0 = 0, pa1 = 0, pe0 = 0, pe1 = 0, pRr = 0, pRe = 0;  // centre row k-1
    float cprev = 0.0f, dprev = 0.0f;
    for (int k = 0; k <= nl; ++k) {
        // centre row k (needs k-1, k, k+1)
        float a0 = 0, a1 = 0, e0 = 0, e1 = 0, Rr = 0, Re = 0;
        if (k < nl) {
            bool hm = k >= 1, hp = k + 1 <= nl - 1;
            jac_center_row(g, k, hm ? &cm : nullptr, c0, hp ? &cp : nullptr, dtg, a0, a1, e0, e1);
            // R = dtg * T_imp(U) for rho and rho e_tot (U = U~ in the first iteration)
            float Fm = 0.0f, Fp = 0.0f, Hm = 0.0f, Hp = 0.0f;
            if (hm) {
                float fu = u3_at(g, u3, k, col) / (g.Jvf[k] * g.Jvf[k]);
                Fm = 0.5f * (cm.rho * cm.J + c0.rho * c0.J) * fu;
                Hm = Fm * (0.5f * (cm.h + c0.h));
            }
            if (hp) {
                float fu = u3_at(g, u3, k + 1, col) / (g.Jvf[k + 1] * g.Jvf[k + 1]);
                Fp = 0.5f * (c0.rho * c0.J + cp.rho * cp.J) * fu;
                Hp = Fp * (0.5f * (c0.h + cp.h));
            }
            Rr = dtg * (-(Fp - Fm) / c0.J);
            Re = dtg * (-(Hp - Hm) / c0.J);
            int64_t q = (int64_t)k * nc + col;
            sa0[q] = a0;
            sa1[q] = a1;
            se0[q] = e0;
            se1[q] = e1;
            sRr[q] = Rr;
            sRe[q] = Re;
        }
        // face row k (needs centre rows k-1 and k and centres k-1, k)
        float L = 0.0f, Dg, Up = 0.0f, r = 0.0f;
        if (k >= 1 && k <= nl - 1) {
            float u3m = u3_at(g, u3, k - 1, col), u3f = u3_at(g, u3, k, col), u3p = u3_at(g, u3, k + 1, col);
            FaceRow fr = jac_face_row(g, n, k, col, cm, c0, u3m, u3f, u3p, dtg);
            float Ru3 = dtg * (pgf_u3(g, n, k, cm, c0) - beta_w(g, n, k, col) * u3f);
            L = fr.lo + fr.r0 * pa0 + fr.e0 * pe0;
            Dg = fr.d + fr.r0 * pa1 + fr.e0 * pe1 + fr.r1 * a0 + fr.e1 * e0;
            Up = fr.up + fr.r1 * a1 + fr.e1 * e1;
            r = Ru3 + fr.r0 * pRr + fr.r1 * Rr + fr.e0 * pRe + fr.e1 * Re;
        } else {
            Dg = -dtg * beta_w(g, n, k, col) - 1.0f;
        }
        float m = Dg - L * cprev;
        float cc = Up / m;
        float dd = (r - L * dprev) / m;
        int64_t qf = (int64_t)k * nc + col;
        cpr[qf] = cc;
        dpr[qf] = dd;
        cprev = cc;
        dprev = dd;
        pa0 = a0;
        pa1 = a1;
        pe0 = e0;
        pe1 = e1;
        pRr = Rr;
        pRe = Re;
        // slide the centre window
        cm = c0;
        c0 = cp;
        if (k + 2 <= nl - 1) cp = cv_load(g, n, U, k + 2, col);
    }
    // ---- pass C: back substitution, U <- U - dU; F_I partials
    float xn = 0.0f;
    for (int f = nl; f >= 0; --f) {
        int64_t q = (int64_t)f * nc + col;
        float x = dpr[q] - cpr[q] * xn;
        float *u3w = U + 4 * nC;
        if (f >= 1 && f <= nl - 1) {
            u3w[q] -= x;
            FI[2 * nC + q] = -x / dtg;
        } else {
            u3w[q] = 0.0f;
            FI[2 * nC + q] = 0.0f;
        }
        if (f < nl) {
            float dr = sa0[q] * x + sa1[q] * xn - sRr[q];
            float de = se0[q] * x + se1[q] * xn - sRe[q];
            U[q] -= dr;
            U[3 * nC + q] -= de;
            FI[q] = -dr / dtg;
            sRe[q] = de;  // keep d(rho e) for pass D
        }
        xn = x;
    }
    // ---- pass D: post-Newton upwind correction (T_post_imp!) at the new U
    if (n.upw != 0) {
        // window of h and rho J over centres f-2 .. f+1
        float hw[4], rw[4];
        for (int t = 0; t < 4; ++t) {
            hw[t] = 0.0f;
            rw[t] = 0.0f;
        }
        // preload centres 0 and 1 into slots 2, 3 (face 0 is a boundary)
        // h is taken relative to its value in the lowest cell: both the upwind
        // and the central product shift by the same u^3 h_ref, which cancels in
        // their difference, and the float32 differences stay accurate
        float href;
        {
            CV c = cv_load(g, n, U, 0, col);
            href = c.h;
            hw[2] = 0.0f;
            rw[2] = c.rho * c.J;
            if (nl > 1) {
                CV d1 = cv_load(g, n, U, 1, col);
                hw[3] = d1.h - href;
                rw[3] = d1.rho * d1.J;
            }
        }
        float dFprev = 0.0f;
        for (int f = 1; f <= nl; ++f) {
            // shift: slots hold centres f-2, f-1, f, f+1
            hw[0] = hw[1];
            hw[1] = hw[2];
            hw[2] = hw[3];
            rw[0] = rw[1];
            rw[1] = rw[2];
            rw[2] = rw[3];
            if (f + 1 <= nl - 1) {
                CV c = cv_load(g, n, U, f + 1, col);
                hw[3] = c.h - href;
                rw[3] = c.rho * c.J;
            } else {
                hw[3] = 0.0f;
                rw[3] = 0.0f;
            }
            float dF = 0.0f;
            if (f <= nl - 1) {
                float v = u3_at(g, U + 4 * nC, f, col) / (g.Jvf[f] * g.Jvf[f]);
                float rJ = 0.5f * (rw[1] + rw[2]);
                float up = upwind_product(n.upw, f, nl, v, n.dt, hw[0], hw[1], hw[2], hw[3]);
                float ce = upwind_product(0, f, nl, v, n.dt, 0.0f, hw[1], hw[2], 0.0f);
                dF = rJ * up - rJ * ce;
            }
            // centre f-1 is complete
            int k = f - 1;
            int64_t q = (int64_t)k * nc + col;
            float J = g.Jh[col] * g.sc[k] * g.sc[k] * g.Jvc[k];
            float corr = -(dF - dFprev) / J;
            U[3 * nC + q] += dtg * corr;
            FI[nC + q] = -sRe[q] / dtg + corr;
            dFprev = dF;
        }
    } else {
        for (int k = 0; k < nl; ++k) {
            int64_t q = (int64_t)k * nc + col;
            FI[nC + q] = -sRe[q] / dtg;
        }
    }
}

// Final combination out = u + dt sum_j b_j (F_E,j + F_I,j) (the DSS and the
// impenetrability filter of the end of the step follow in k_dss_gather).
__global__ void __launch_bounds__(256) k_combine(DevGeom g, const float *u, float *out, FinalSrc s) {
    // four consecutive values per thread (every plane size is a multiple of 4
    // when ncol is; otherwise the scalar tail path is taken)
    const int64_t S = 4 * g.nC + g.nF;
    const int64_t i0 = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) * 4;
    if (i0 >= S) return;
    const bool vec = (g.ncol & 3) == 0 && i0 + 3 < S;
    const int comp = i0 < 4 * g.nC ? (int)(i0 / g.nC) : 4;
    const int64_t r = i0 - (int64_t)comp * g.nC;
    const bool has_fi = s.fi_full || comp == 0 || comp == 3 || comp == 4;
    const int ci = s.fi_full ? comp : (comp == 0 ? 0 : (comp == 3 ? 1 : 2));
    if (vec) {
        float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
        for (int t = 0; t < s.nfe; ++t) {
            const float4 v = *reinterpret_cast<const float4 *>(s.fe[t] + i0);
            acc.x = fmaf(s.be[t], v.x, acc.x);
            acc.y = fmaf(s.be[t], v.y, acc.y);
            acc.z = fmaf(s.be[t], v.z, acc.z);
            acc.w = fmaf(s.be[t], v.w, acc.w);
        }
        if (has_fi)
            for (int t = 0; t < s.nfi; ++t) {
                const float4 v = *reinterpret_cast<const float4 *>(s.fi[t] + ci * g.nC + r);
                acc.x = fmaf(s.bi[t], v.x, acc.x);
                acc.y = fmaf(s.bi[t], v.y, acc.y);
                acc.z = fmaf(s.bi[t], v.z, acc.z);
                acc.w = fmaf(s.bi[t], v.w, acc.w);
            }
        const float4 u4 = *reinterpret_cast<const float4 *>(u + i0);
        float4 o;
        o.x = u4.x + s.dt * acc.x;
        o.y = u4.y + s.dt * acc.y;
        o.z = u4.z + s.dt * acc.z;
        o.w = u4.w + s.dt * acc.w;
        *reinterpret_cast<float4 *>(out + i0) = o;
        return;
    }
    for (int64_t idx = i0; idx < i0 + 4 && idx < S; ++idx) {
        const int cp = idx < 4 * g.nC ? (int)(idx / g.nC) : 4;
        const int64_t rr = idx - (int64_t)cp * g.nC;
        float acc = 0.0f;
        for (int t = 0; t < s.nfe; ++t) acc = fmaf(s.be[t], s.fe[t][idx], acc);
        if (s.fi_full || cp == 0 || cp == 3 || cp == 4) {
            const int cj = s.fi_full ? cp : (cp == 0 ? 0 : (cp == 3 ? 1 : 2));
            for (int t = 0; t < s.nfi; ++t) acc = fmaf(s.bi[t], s.fi[t][cj * g.nC + rr], acc);
        }
        out[idx] = u[idx] + s.dt * acc;
    }
}
