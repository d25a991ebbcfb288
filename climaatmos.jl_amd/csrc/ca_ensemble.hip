// ca_ensemble.hip -- batched single-column ensembles of uneven vertical sizes
// with per-member parameters, stepped by one persistent launch.
//
// Each member is an independent column (ColumnGrid, simulation/grids.jl:110-129)
// with its own face heights, Rayleigh coefficient alpha_w
// (rayleigh_sponge.jl:46-57) and vertical diffusivity D0
// (vertical_diffusion_boundary_layer.jl:63-103, H = Inf).  A step is the same
// IMEX ARS343 step with one Newton iteration as ca_step on a column grid:
// explicit tendency, three implicit Newton stages with the manual Jacobian,
// the Schur complement on u_3 and the post-Newton upwind correction, and the
// final combination with the impenetrability filter.
//
// Mapping: one workgroup of LPM lanes per member (LPM = 64, 128 or 256, the
// autotuner picks it), levels strided over the lanes.  The member's state,
// its four explicit and three implicit stage tendencies and the stage state
// stay in LDS for all nsteps steps; HBM is read once and written once per
// launch.  The u_3 Schur system (faces 1 .. nlev - 1; the boundary faces
// carry u_3 = 0) is solved by parallel cyclic reduction in float64 through
// LDS.  The member order (as given, or longest first) is a second knob.
#include "ca_kernels.h"

#define CA_ENS_MAXL 128             // centre levels per member, at most
#define CA_ENS_NF (CA_ENS_MAXL + 1)  // faces

struct EnsTab {
    float aE[4][4], aI[4][4], b[4];
};

namespace ens {

// Newton-stage intermediates of one column (LDS)
template <int NF>
struct NwT {
    float c[11][NF];       // rho, K, T, p, Pi, ct1, ct2, h, J, iJ, rJ
    double cd[3][NF];      // Td, pd, Pid
    float row[6][NF];      // a0, a1, e0, e1, Rr, Re
    double sys[2][4][NF];  // PCR ping-pong: L, D, U, r
    float x[NF];           // du_3
    float dF[NF];          // upwind - central flux difference
    float hr[NF];          // h relative to the lowest cell, new state
    float rJ[NF];          // rho J, new state
    float ve[NF];          // rho e_tot after the Newton update, before the correction
};

// LDS of one member.  The scratch union holds the explicit-tendency or the
// Newton-stage intermediates.
template <int NF>
struct SharedT {
    static constexpr int NFv = NF;
    float u[5][NF];      // state at the start of the step (4 centre planes, u_3)
    float fe[4][5][NF];  // explicit tendencies of the four stages
    float fi[3][3][NF];  // implicit tendencies of stages 1..3: rho, rho e_tot, u_3
    float U[5][NF];      // stage state
    union {
        struct {
            float c[10][NF];  // rhoJ, Kh, Kv, u1, u2, rho, sd, irhoJ, d1, d2
            float f[5][NF];   // G1, G2, Fs, Fu0, Fu1
        } ex;
        NwT<NF> nw;
    };
};

// The part of SharedT one implicit stage uses (newton_stage alone)
template <int NF>
struct StageShared {
    static constexpr int NFv = NF;
    float fi[1][3][NF];
    float U[5][NF];
    NwT<NF> nw;
};

enum { EC_RHOJ, EC_KH, EC_KV, EC_U1, EC_U2, EC_RHO, EC_SD, EC_IRHOJ, EC_D1, EC_D2 };
enum { NC_RHO, NC_K, NC_T, NC_P, NC_PI, NC_CT1, NC_CT2, NC_H, NC_J, NC_IJ, NC_RJ };

__device__ __forceinline__ void sync() { __syncthreads(); }

// CV of centre k from the Newton scratch
template <class SH>
__device__ __forceinline__ CV cv_from(const SH &S, int k) {
    CV c;
    c.rho = S.nw.c[NC_RHO][k];
    c.K = S.nw.c[NC_K][k];
    c.T = S.nw.c[NC_T][k];
    c.p = S.nw.c[NC_P][k];
    c.Pi = S.nw.c[NC_PI][k];
    c.ct1 = S.nw.c[NC_CT1][k];
    c.ct2 = S.nw.c[NC_CT2][k];
    c.h = S.nw.c[NC_H][k];
    c.J = S.nw.c[NC_J][k];
    c.iJ = S.nw.c[NC_IJ][k];
    c.Td = S.nw.cd[0][k];
    c.pd = S.nw.cd[1][k];
    c.Pid = S.nw.cd[2][k];
    c.u1 = c.u2 = c.rhoe = 0.0f;
    return c;
}

// remaining_tendency! of a column (no horizontal terms; shallow f-plane
// branch of advection.jl:262-283, Rayleigh u_h sponge, K-theory vertical
// diffusion): Y (LDS planes) -> T (LDS planes).  Same arithmetic as the
// column path of k_explicit.
template <int LPM, int NF>
__device__ void explicit_tendency(SharedT<NF> &S, const DevGeom &g, const DevNum &n, const float (*Y)[NF],
                                  float (*T)[NF], int tid) {
    const int nl = g.nlev;
    const float f3n = g.f3n[0], f12a = g.f12a[0], f12b = g.f12b[0];
    float Kf = 0.0f;
    if (n.vdiff) {
        const float Kh = fmaxf(g.D0_col[0], 1.1920929e-07f);
        Kf = 1.0f / (0.5f * (1.0f / Kh + 1.0f / Kh));
    }
    for (int k = tid; k < nl; k += LPM) {
        const float rho = Y[0][k], u1 = Y[1][k], u2 = Y[2][k], rhoe = Y[3][k];
        const float u3a = k == 0 ? 0.0f : Y[4][k], u3b = k + 1 >= nl ? 0.0f : Y[4][k + 1];
        const float J = g.Jh[0] * g.Jvc[k];
        float ct1, ct2;
        ca_contra(g, 0, g.is2c[k], u1, u2, ct1, ct2);
        const float Khor = 0.5f * (u1 * ct1 + u2 * ct2);
        const float Kver = 0.25f * (u3a * g.g33f[k] * u3a + u3b * g.g33f[k + 1] * u3b);
        const float K = Khor + Kver;
        const double e_tot_d = (double)rhoe * ca_drcp((double)rho);
        const double Td = fmax(n.dT_0 + (e_tot_d - (double)K - g.Phic_d[k] + n.dR_d * n.dT_0) * n.dicv_d, n.dTmin);
        const float Tf = (float)Td;
        float d1 = 0.0f, d2 = 0.0f;
        const float a3 = f3n * g.iJvc[k];
        d1 -= J * (0.0f - a3 * ct2);
        d2 -= J * (a3 * ct1 - 0.0f);
        if (n.rayleigh && n.alpha_uh != 0.0f) {
            const float beta = n.alpha_uh * g.zeta_c[k];
            d1 += -beta * u1;
            d2 += -beta * u2;
        }
        S.ex.c[EC_RHOJ][k] = rho * J;
        S.ex.c[EC_KH][k] = Khor;
        S.ex.c[EC_KV][k] = Kver;
        S.ex.c[EC_U1][k] = ct1;
        S.ex.c[EC_U2][k] = ct2;
        S.ex.c[EC_RHO][k] = rho;
        S.ex.c[EC_SD][k] = n.cp_d * (Tf - n.T_0) + g.Phic[k];
        S.ex.c[EC_IRHOJ][k] = ca_rcp(rho * J);
        S.ex.c[EC_D1][k] = d1;
        S.ex.c[EC_D2][k] = d2;
    }
    sync();
    for (int f = tid; f <= nl; f += LPM) {
        float G1 = 0.0f, G2 = 0.0f, du3 = 0.0f, Fs = 0.0f, Fu0 = 0.0f, Fu1 = 0.0f;
        if (f >= 1 && f <= nl - 1) {
            const int a = f - 1, b = f;
            const float Jf = g.Jh[0] * g.sf[f] * g.sf[f] * g.Jvf[f], iJf = g.iJh[0] * g.isJf[f];
            const float u3f = Y[4][f];
            const float fu3 = u3f * g.g33f[f];
            // u_1, u_2 here are the covariant components (column metric: identity)
            const float wa = -(Y[2][b] - Y[2][a]) * iJf;
            const float wb = (Y[1][b] - Y[1][a]) * iJf;
            const float A1 = f12a * g.isf[f] + wa, A2 = f12b * g.isf[f] + wb;
            const float rJf = 0.5f * (S.ex.c[EC_RHOJ][a] + S.ex.c[EC_RHOJ][b]);
            const float b3 = rJf * fu3;
            G1 = Jf * (A2 * b3);
            G2 = Jf * (0.0f - A1 * b3);
            const float i1 = 0.5f * (S.ex.c[EC_U1][a] + S.ex.c[EC_U1][b]);
            const float i2 = 0.5f * (S.ex.c[EC_U2][a] + S.ex.c[EC_U2][b]);
            du3 = -(Jf * (A1 * i2 - A2 * i1) +
                    ((S.ex.c[EC_KH][b] - S.ex.c[EC_KH][a]) + (S.ex.c[EC_KV][b] - S.ex.c[EC_KV][a])));
            if (n.vdiff) {
                const float rK = (0.5f * (S.ex.c[EC_RHO][a] + S.ex.c[EC_RHO][b])) * Kf;
                const float iJv = g.iJvf[f];
                Fs = Jf * (iJv * iJv * (-(rK * (S.ex.c[EC_SD][b] - S.ex.c[EC_SD][a]))));
                const float fac = Jf * iJv * (-2.0f * rK) * 0.5f;
                Fu0 = fac * (Y[1][b] - Y[1][a]) * iJv;
                Fu1 = fac * (Y[2][b] - Y[2][a]) * iJv;
            }
        }
        S.ex.f[0][f] = G1;
        S.ex.f[1][f] = G2;
        S.ex.f[2][f] = Fs;
        S.ex.f[3][f] = Fu0;
        S.ex.f[4][f] = Fu1;
        T[4][f] = du3;
    }
    sync();
    for (int k = tid; k < nl; k += LPM) {
        const float irJ = S.ex.c[EC_IRHOJ][k];
        float d1 = S.ex.c[EC_D1][k] - (0.5f * (S.ex.f[0][k] + S.ex.f[0][k + 1])) * irJ;
        float d2 = S.ex.c[EC_D2][k] - (0.5f * (S.ex.f[1][k] + S.ex.f[1][k + 1])) * irJ;
        float de = 0.0f;
        if (n.vdiff) {
            const float iJc = g.iJh[0] * g.isJc[k];
            de -= (S.ex.f[2][k + 1] - S.ex.f[2][k]) * iJc;
            const float e0 = (S.ex.f[3][k + 1] - S.ex.f[3][k]) * iJc, e1 = (S.ex.f[4][k + 1] - S.ex.f[4][k]) * iJc;
            const float irho = ca_rcp(S.ex.c[EC_RHO][k]);
            d1 -= e0 * irho;
            d2 -= e1 * irho;
        }
        T[0][k] = 0.0f;
        T[1][k] = d1;
        T[2][k] = d2;
        T[3][k] = de;
    }
    sync();
}

// Thermodynamic state of the stage state U~ in S.U, centre by centre, into
// S.nw.c and S.nw.cd.
template <int LPM, class SH>
__device__ __forceinline__ void stage_thermo(SH &S, const DevGeom &g, const DevNum &n, int tid) {
    const int nl = g.nlev;
    for (int k = tid; k < nl; k += LPM) {
        const float u3a = k == 0 ? 0.0f : S.U[4][k], u3b = k + 1 >= nl ? 0.0f : S.U[4][k + 1];
        const CV c = cv_make(g, n, k, 0, S.U[0][k], S.U[1][k], S.U[2][k], S.U[3][k], u3a, u3b);
        S.nw.c[NC_RHO][k] = c.rho;
        S.nw.c[NC_K][k] = c.K;
        S.nw.c[NC_T][k] = c.T;
        S.nw.c[NC_P][k] = c.p;
        S.nw.c[NC_PI][k] = c.Pi;
        S.nw.c[NC_CT1][k] = c.ct1;
        S.nw.c[NC_CT2][k] = c.ct2;
        S.nw.c[NC_H][k] = c.h;
        S.nw.c[NC_J][k] = c.J;
        S.nw.c[NC_IJ][k] = c.iJ;
        S.nw.c[NC_RJ][k] = c.rho * c.J;
        S.nw.cd[0][k] = c.Td;
        S.nw.cd[1][k] = c.pd;
        S.nw.cd[2][k] = c.Pid;
    }
}

// Centre rows of the Newton system at U~: A(rho, u3), A(rho e, u3) on faces
// k, k+1 and R = dtg T_imp, into S.nw.row (reads S.nw.c).
template <int LPM, class SH>
__device__ __forceinline__ void stage_rows(SH &S, const DevGeom &g, float dtg, int tid) {
    const int nl = g.nlev;
    for (int k = tid; k < nl; k += LPM) {
        const float rJ = S.nw.c[NC_RJ][k], h = S.nw.c[NC_H][k], iJ = S.nw.c[NC_IJ][k];
        float a0 = 0.0f, a1 = 0.0f, e0 = 0.0f, e1 = 0.0f, Fm = 0.0f, Fp = 0.0f, Hm = 0.0f, Hp = 0.0f;
        if (k >= 1) {
            const float rJf = 0.5f * (S.nw.c[NC_RJ][k - 1] + rJ);
            const float g33 = g.g33f[k];
            const float hf = 0.5f * (S.nw.c[NC_H][k - 1] + h);
            a0 = dtg * (rJf * iJ) * g33;
            e0 = a0 * hf;
            Fm = rJf * (S.U[4][k] * g33);
            Hm = Fm * hf;
        }
        if (k + 1 <= nl - 1) {
            const float rJf = 0.5f * (rJ + S.nw.c[NC_RJ][k + 1]);
            const float g33 = g.g33f[k + 1];
            const float hf = 0.5f * (h + S.nw.c[NC_H][k + 1]);
            a1 = dtg * (-rJf * iJ) * g33;
            e1 = a1 * hf;
            Fp = rJf * (S.U[4][k + 1] * g33);
            Hp = Fp * hf;
        }
        S.nw.row[0][k] = a0;
        S.nw.row[1][k] = a1;
        S.nw.row[2][k] = e0;
        S.nw.row[3][k] = e1;
        S.nw.row[4][k] = dtg * (-(Fp - Fm) * iJ);
        S.nw.row[5][k] = dtg * (-(Hp - Hm) * iJ);
    }
}

// Enthalpy relative to the lowest cell and rho J of the post-Newton state in
// S.U, the inputs of the upwind correction, into S.nw.hr and S.nw.rJ.
template <int LPM, class SH>
__device__ __forceinline__ void stage_enthalpy(SH &S, const DevGeom &g, const DevNum &n, int tid) {
    const int nl = g.nlev;
    for (int k = tid; k < nl; k += LPM) {
        const float u3a = k == 0 ? 0.0f : S.U[4][k], u3b = k + 1 >= nl ? 0.0f : S.U[4][k + 1];
        S.nw.hr[k] = cv_enthalpy(g, n, k, 0, S.U[0][k], S.U[1][k], S.U[2][k], S.U[3][k], u3a, u3b);
        S.nw.rJ[k] = S.U[0][k] * S.nw.c[NC_J][k];
    }
}

// One implicit stage: S.U holds U~ on entry and U on exit; fi receives
// F_I = (U - U~) / dtg for rho, rho e_tot and u_3.  Same algorithm as
// k_newton_pcr (one Newton iteration from U~, manual Jacobian, Schur
// complement on u_3, post-Newton upwind correction).
template <int LPM, class SH>
__device__ void newton_stage(SH &S, const DevGeom &g, const DevNum &n, float dtg, float (*fi)[SH::NFv],
                             int tid) {
    const int nl = g.nlev;
    const float idtg = 1.0f / dtg;
    // centres: thermodynamic state of U~
    stage_thermo<LPM>(S, g, n, tid);
    sync();
    // centre rows: A(rho, u3), A(rho e, u3) on faces k, k+1 and R = dtg T_imp
    stage_rows<LPM>(S, g, dtg, tid);
    sync();
    // face rows of the Schur system, faces 0 .. nl - 1 (face nl: u_3 = 0, x = 0)
    for (int f = tid; f < nl; f += LPM) {
        double L = 0.0, D, U = 0.0, r = 0.0;
        if (f >= 1) {
            const CV a = cv_from(S, f - 1), c = cv_from(S, f);
            const float u3m = f - 1 >= 1 ? S.U[4][f - 1] : 0.0f;
            const float u3 = S.U[4][f];
            const float u3p = f + 1 <= nl - 1 ? S.U[4][f + 1] : 0.0f;
            const FaceRow fr = jac_face_row(g, n, f, 0, a, c, u3m, u3, u3p, dtg);
            const float Ru3 = dtg * (pgf_u3(g, n, f, a, c) - beta_w(g, n, f, 0) * u3);
            const float pa0 = S.nw.row[0][f - 1], pa1 = S.nw.row[1][f - 1];
            const float pe0 = S.nw.row[2][f - 1], pe1 = S.nw.row[3][f - 1];
            const float a0 = S.nw.row[0][f], e0 = S.nw.row[2][f], a1 = S.nw.row[1][f], e1 = S.nw.row[3][f];
            L = fr.lo + fr.r0 * pa0 + fr.e0 * pe0;
            D = fr.d + fr.r0 * pa1 + fr.e0 * pe1 + fr.r1 * a0 + fr.e1 * e0;
            U = f + 1 <= nl - 1 ? fr.up + fr.r1 * a1 + fr.e1 * e1 : 0.0;
            r = Ru3 + fr.r0 * S.nw.row[4][f - 1] + fr.r1 * S.nw.row[4][f] + fr.e0 * S.nw.row[5][f - 1] +
                fr.e1 * S.nw.row[5][f];
        } else {
            D = -dtg * beta_w(g, n, 0, 0) - 1.0f;
        }
        S.nw.sys[0][0][f] = L;
        S.nw.sys[0][1][f] = D;
        S.nw.sys[0][2][f] = U;
        S.nw.sys[0][3][f] = r;
    }
    sync();
    // parallel cyclic reduction in float64 over the nl face rows: step s
    // eliminates the couplings to rows f -+ s (absent beyond the ends)
    int src = 0;
    for (int s = 1; s < nl; s <<= 1) {
        for (int f = tid; f < nl; f += LPM) {
            const double L = S.nw.sys[src][0][f], D = S.nw.sys[src][1][f], U = S.nw.sys[src][2][f],
                         r = S.nw.sys[src][3][f];
            double nL = 0.0, nU = 0.0, nD = D, nr = r;
            if (f - s >= 0) {
                const double al = -L * ca_drcp(S.nw.sys[src][1][f - s]);
                nL = al * S.nw.sys[src][0][f - s];
                nD += al * S.nw.sys[src][2][f - s];
                nr += al * S.nw.sys[src][3][f - s];
            }
            if (f + s < nl) {
                const double be = -U * ca_drcp(S.nw.sys[src][1][f + s]);
                nU = be * S.nw.sys[src][2][f + s];
                nD += be * S.nw.sys[src][0][f + s];
                nr += be * S.nw.sys[src][3][f + s];
            }
            S.nw.sys[src ^ 1][0][f] = nL;
            S.nw.sys[src ^ 1][1][f] = nD;
            S.nw.sys[src ^ 1][2][f] = nU;
            S.nw.sys[src ^ 1][3][f] = nr;
        }
        sync();
        src ^= 1;
    }
    for (int f = tid; f <= nl; f += LPM)
        S.nw.x[f] = f < nl ? (float)(S.nw.sys[src][3][f] * ca_drcp(S.nw.sys[src][1][f])) : 0.0f;
    sync();
    // Newton update and F_I (faces first: u_3; then centres: rho, rho e_tot)
    for (int f = tid; f <= nl; f += LPM) {
        const bool intF = f >= 1 && f <= nl - 1;
        const float x = S.nw.x[f];
        S.U[4][f] = intF ? S.U[4][f] - x : 0.0f;
        fi[2][f] = intF ? -x * idtg : 0.0f;
    }
    for (int k = tid; k < nl; k += LPM) {
        const float x0 = S.nw.x[k], x1 = S.nw.x[k + 1];
        const float dr = S.nw.row[0][k] * x0 + S.nw.row[1][k] * x1 - S.nw.row[4][k];
        const float de = S.nw.row[2][k] * x0 + S.nw.row[3][k] * x1 - S.nw.row[5][k];
        S.U[0][k] -= dr;
        S.U[3][k] -= de;
        S.nw.ve[k] = S.U[3][k];
        fi[0][k] = -dr * idtg;
        fi[1][k] = -de * idtg;
    }
    sync();
    // post-Newton upwind correction at the new state (implicit_tendency.jl:322-339)
    if (n.upw != 0) {
        stage_enthalpy<LPM>(S, g, n, tid);
        sync();
        const float href = S.nw.hr[0];
        for (int f = tid; f <= nl; f += LPM) {
            float dF = 0.0f;
            if (f >= 1 && f <= nl - 1) {
                const float v = S.U[4][f] * g.g33f[f];
                const float rJf = 0.5f * (S.nw.rJ[f - 1] + S.nw.rJ[f]);
                const float h2 = f >= 2 ? S.nw.hr[f - 2] - href : 0.0f;
                const float h1 = S.nw.hr[f - 1] - href, h3 = S.nw.hr[f] - href;
                const float h4 = f + 1 <= nl - 1 ? S.nw.hr[f + 1] - href : 0.0f;
                const float upw = upwind_product(n.upw, f, nl, v, n.dt, h2, h1, h3, h4);
                const float cen = upwind_product(0, f, nl, v, n.dt, 0.0f, h1, h3, 0.0f);
                dF = rJf * upw - rJf * cen;
            }
            S.nw.dF[f] = dF;
        }
        sync();
        for (int k = tid; k < nl; k += LPM) {
            const float iJ = S.nw.c[NC_IJ][k];  // J does not depend on the state
            const float corr = -(S.nw.dF[k + 1] - S.nw.dF[k]) * iJ;
            S.U[3][k] += dtg * corr;
            fi[1][k] += corr;
        }
        sync();
    }
}

// One ARS343 step of the member in S.u (in place).  With rUt / rUs the
// stage states U~_i and U_i are also written there (the adjoint's recompute).
template <int LPM, int NF>
__device__ void step_once(SharedT<NF> &S, const DevGeom &g, const DevNum &n, const EnsTab &tab, float dt, int tid,
                          float (*rUt)[5][NF] = nullptr, float (*rUs)[5][NF] = nullptr) {
    const int nl = g.nlev;
    explicit_tendency<LPM>(S, g, n, S.u, S.fe[0], tid);
    for (int i = 1; i < 4; ++i) {
        // stage state U~_i = u + dt sum_j (aE F_E_j + aI F_I_j)
        for (int q = tid; q < 5 * NF; q += LPM) {
            const int p = q / NF, k = q - p * NF;
            if (k > (p < 4 ? nl - 1 : nl)) continue;
            float s = 0.0f;
            for (int j = 0; j < i; ++j) {
                s = fmaf(tab.aE[i][j], S.fe[j][p][k], s);
                if (j >= 1 && (p == 0 || p == 3 || p == 4))
                    s = fmaf(tab.aI[i][j], S.fi[j - 1][p == 0 ? 0 : (p == 3 ? 1 : 2)][k], s);
            }
            float v = fmaf(dt, s, S.u[p][k]);
            if (p == 4 && (k == 0 || k == nl)) v = 0.0f;
            S.U[p][k] = v;
            if (rUt) rUt[i - 1][p][k] = v;
        }
        sync();
        newton_stage<LPM>(S, g, n, dt * tab.aI[i][i], S.fi[i - 1], tid);
        if (rUs) {
            for (int q = tid; q < 5 * NF; q += LPM) {
                const int p = q / NF, k = q - p * NF;
                if (k > (p < 4 ? nl - 1 : nl)) continue;
                rUs[i - 1][p][k] = S.U[p][k];
            }
        }
        explicit_tendency<LPM>(S, g, n, S.U, S.fe[i], tid);
    }
    // u <- u + dt sum_j b_j (F_E_j + F_I_j), filter u_3
    for (int q = tid; q < 5 * NF; q += LPM) {
        const int p = q / NF, k = q - p * NF;
        if (k > (p < 4 ? nl - 1 : nl)) continue;
        float s = 0.0f;
        for (int j = 0; j < 4; ++j) {
            if (tab.b[j] == 0.0f) continue;
            float v = S.fe[j][p][k];
            if (j >= 1 && (p == 0 || p == 3 || p == 4)) v += S.fi[j - 1][p == 0 ? 0 : (p == 3 ? 1 : 2)][k];
            s = fmaf(tab.b[j], v, s);
        }
        float v = fmaf(dt, s, S.u[p][k]);
        if (p == 4 && (k == 0 || k == nl)) v = 0.0f;
        S.u[p][k] = v;
    }
    sync();
}

}  // namespace ens

// Persistent ensemble step: workgroup b steps member order[b] nsteps times.
template <int LPM, int NF>
__global__ void __launch_bounds__(LPM) k_ensemble(const DevGeom *__restrict__ mg, DevNum n, const int *__restrict__ order,
                                                  const int64_t *__restrict__ off, float *__restrict__ Y, EnsTab tab,
                                                  float dt, int nsteps) {
    using namespace ens;
    __shared__ SharedT<NF> S;
    const int m = order[blockIdx.x];
    const DevGeom &g = mg[m];
    const int nl = g.nlev;
    const int tid = threadIdx.x;
    float *Ym = Y + off[m];
    // load, with the impenetrability filter on u_3
    for (int k = tid; k < nl; k += LPM)
        for (int p = 0; p < 4; ++p) S.u[p][k] = Ym[p * nl + k];
    for (int f = tid; f <= nl; f += LPM) S.u[4][f] = (f == 0 || f == nl) ? 0.0f : Ym[4 * nl + f];
    sync();
    for (int st = 0; st < nsteps; ++st) step_once<LPM>(S, g, n, tab, dt, tid);
    for (int k = tid; k < nl; k += LPM)
        for (int p = 0; p < 4; ++p) Ym[p * nl + k] = S.u[p][k];
    for (int f = tid; f <= nl; f += LPM) Ym[4 * nl + f] = S.u[4][f];
}

#define CA_ENS_INST(LPM, NF)                                                                                    \
    template __global__ void k_ensemble<LPM, NF>(const DevGeom *, DevNum, const int *, const int64_t *, float *, EnsTab, \
                                                 float, int);
CA_ENS_INST(64, 33)
CA_ENS_INST(128, 33)
CA_ENS_INST(64, 65)
CA_ENS_INST(128, 65)
CA_ENS_INST(256, 65)
CA_ENS_INST(64, 129)
CA_ENS_INST(128, 129)
CA_ENS_INST(256, 129)
#undef CA_ENS_INST
