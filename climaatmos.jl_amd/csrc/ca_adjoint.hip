// ca_adjoint.hip -- reverse mode (vector-Jacobian product) of the column
// ARS343 step over many steps, with step checkpointing.
//
// No reference counterpart: the reference has forward-mode AD for its
// Jacobians (src/prognostic_equations/implicit/autodiff_utils.jl) and
// gradient-free calibration (calibration/).  Given the cotangent of the
// state after nsteps steps, k_ensemble_adjoint returns the cotangent of the
// initial state and of each member's vertical diffusivity D0 and Rayleigh
// coefficient alpha_w, for the same step k_ensemble takes:
//
//   forward:  u -> FE_0 = T_E(u); for i = 1..3: U~_i = u + dt sum_j<i (aE FE_j + aI FI_j),
//             U_i = Newton_i(U~_i) (+ upwind correction), FI_i = (U_i - U~_i) / dtg_i,
//             FE_i = T_E(U_i);  u' = u + dt sum_j b_j (FE_j + FI_j)
//   reverse:  the same sweep backwards, stage 3 to 1, with the transposes of
//             T_E (explicit tendency), of the upwind correction, and of the
//             Newton update; the latter holds the transposed Schur solve
//             (S^T lambda = x_bar, a tridiagonal system solved by parallel
//             cyclic reduction in float64) and the derivatives of the
//             Jacobian entries, which are taken in forward mode over the 11
//             values a face row reads (dual numbers, float64).
//
// Checkpointing: the forward pass stores the state every `every` steps; the
// reverse pass recomputes each segment's step states from its checkpoint and
// each step's stage states from its start state.  The non-smooth points
// (the T_min clamp, the van Leer limiter branches) take the derivative of the
// branch the forward pass took.
#include "ca_kernels.h"

namespace adj {


// ------------------------------------------------------- face row reverse
// The 11 values a face row of the Newton stage reads: rho, u1, u2, rho e_tot
// of the centres below (a) and above (b) the face, and u_3 on the faces
// f - 1, f, f + 1.  Its outputs, as in jac_face_row and pgf_u3: lo, d, up,
// r0, r1, e0, e1 of the Schur row and Ru3 = dtg (pgf - beta u3).
constexpr int NP = 11;

// thermodynamic state of one centre in float64 (cv_make's arithmetic)
struct Cen {
    double rho, u1, u2, ct1, ct2, re, e, K, T, p, Pi;
    bool free_T;
};
// Td, pd, Pid: T, p and Pi of the same state as the Newton stage stored them
__device__ __forceinline__ Cen cen_eval(const DevGeom &g, const DevNum &n, int k, double rho, double u1, double u2,
                                        double re, double u3a, double u3b, double Td, double pd, double Pid) {
    Cen c;
    c.rho = rho;
    c.u1 = u1;
    c.u2 = u2;
    c.re = re;
    const double is2 = g.is2c[k];
    c.ct1 = is2 * ((double)g.gu11[0] * u1 + (double)g.gu12[0] * u2);
    c.ct2 = is2 * ((double)g.gu12[0] * u1 + (double)g.gu22[0] * u2);
    c.K = 0.5 * ((u1 * c.ct1 + u2 * c.ct2) + 0.5 * ((double)g.g33f[k] * (u3a * u3a) + (double)g.g33f[k + 1] * (u3b * u3b)));
    c.e = re * ca_drcp(rho);  // as the forward forms e_tot
    c.T = Td;
    c.free_T = Td > n.dTmin;  // the clamp T_min sets Td to exactly T_min
    c.p = pd;
    c.Pi = Pid;
    return c;
}
// reverse of cen_eval: cotangents of T, p (direct) and Pi -> rho, u1, u2, rho e_tot, u3a, u3b
__device__ __forceinline__ void cen_reverse(const DevGeom &g, const DevNum &n, int k, const Cen &c, double u3a,
                                            double u3b, double Tb, double pb, double Pib, double Kb, double *out) {
    pb += Pib * n.dkex * c.Pi * ca_drcp(c.p);
    double rhob = n.dR_d * c.T * pb;
    Tb += n.dR_d * c.rho * pb;
    const double Tub = c.free_T ? Tb : 0.0;
    const double eb = n.dicv_d * Tub;
    Kb -= n.dicv_d * Tub;
    const double irho = ca_drcp(c.rho);
    out[3] += eb * irho;
    rhob -= eb * c.e * irho;
    out[0] += rhob;
    // K = (u . ct) / 2 + (g_a u3a^2 + g_b u3b^2) / 4, ct = is2 G^-1 u (symmetric)
    out[1] += Kb * c.ct1;
    out[2] += Kb * c.ct2;
    out[4] += Kb * 0.5 * (double)g.g33f[k] * u3a;
    out[5] += Kb * 0.5 * (double)g.g33f[k + 1] * u3b;
}

struct FaceVals {
    double r0, r1, e0, e1;
};

// Face row f forward (float64) and its reverse seeded by ob = the
// cotangents of (lo, d, up, r0, r1, e0, e1, Ru3): pb receives the
// cotangents of the 11 inputs.  Returns the row's r0, r1, e0, e1.
// tpa, tpb: (Td, pd, Pid) of the centres below and above, as stored.
__device__ __forceinline__ FaceVals face_vjp(const DevGeom &g, const DevNum &n, int f, float dtgf, const float *prim,
                                             const double *tpa, const double *tpb, const double *ob, double *pb) {
    const double dtg = dtgf;
    const double x8 = prim[8], x9 = prim[9], x10 = prim[10];
    const Cen a = cen_eval(g, n, f - 1, prim[0], prim[1], prim[2], prim[3], x8, x9, tpa[0], tpa[1], tpa[2]);
    const Cen b = cen_eval(g, n, f, prim[4], prim[5], prim[6], prim[7], x9, x10, tpb[0], tpb[1], tpb[2]);
    const double kap = n.dR_d / n.dcv_d, cp = n.dcp_d, dTs = n.dTsr - n.dTmr;
    // ---- forward
    const double rf = 0.5 * (a.rho + b.rho);
    const double pgm = ca_drcp(rf), pgp = -pgm;
    const double dpa = kap * ((n.dT_0 * cp - (double)g.Phic[f - 1]) - a.K);
    const double dpb = kap * ((n.dT_0 * cp - (double)g.Phic[f]) - b.K);
    // reciprocals (v_rcp_f64 and two Newton-Raphson steps) in place of the
    // correctly rounded float64 quotients
    const double iPa = ca_drcp(a.Pi), iPib = ca_drcp(b.Pi), ipa = ca_drcp(a.p);
    const double thva = a.T * iPa, thvb = b.T * iPib;
    const double z = (b.p - a.p) * ipa;
    const double lr = n.dkex * log1p(z);
    const double Em = expm1(lr);
    const double dPi = a.Pi * Em;
    const double thbar = 0.5 * (thva + thvb);
    const double buoy = cp * thbar * dPi * pgm;
    FaceVals v;
    v.r0 = dtg * (pgm * dpa + 0.5 * buoy);
    v.r1 = dtg * (pgp * dpb + 0.5 * buoy);
    v.e0 = dtg * kap * pgm;
    v.e1 = dtg * kap * pgp;
    const double ka = -kap * a.rho, kb = -kap * b.rho;
    const double gm = g.g33f[f - 1], g0 = g.g33f[f], gp = g.g33f[f + 1];
    const double beta = n.rayleigh ? (double)g.aw_col[0] * (double)g.zeta_f[f] : 0.0;
    const double Pa7 = a.Pi * a.Pi * a.Pi * a.Pi * a.Pi * a.Pi * a.Pi;
    const double E7 = expm1(7.0 * lr);
    const double Pb = a.Pi + dPi, iPb = ca_drcp(Pb);
    const double Pb7 = Pb * Pb * Pb * Pb * Pb * Pb * Pb;
    const double tha = (a.T - (n.dTmr + dTs * Pa7)) * iPa;
    const double thb = (b.T - (n.dTmr + dTs * Pb7)) * iPb;
    // ---- reverse
    const double Lb = ob[0], Db = ob[1], Ub = ob[2], r0b = ob[3], r1b = ob[4], e0b = ob[5], e1b = ob[6], qb = ob[7];
    double x8b = 0.0, x9b = 0.0, x10b = 0.0;
    // Ru3 = dtg (pgf - beta u3); pgf = -((dPhi - dPhr) + cp (tha + thb)/2 dPi)
    const double pgfb = dtg * qb;
    x9b -= dtg * beta * qb;
    const double dPhrb = pgfb;
    const double thab = -0.5 * cp * dPi * pgfb, thbb = thab;
    double dPib = -cp * 0.5 * (tha + thb) * pgfb;
    // thb = (T_b - T_r(Pb)) / Pb, tha likewise with Pa
    double Tbb = thbb * iPb, Tab = thab * iPa;
    double Pbb = -thbb * thb * iPb - thbb * dTs * 7.0 * (Pb7 * iPb) * iPb;
    double Pab = -thab * tha * iPa;
    double Pa7b = -dTs * thab * iPa;
    Pab += Pbb;
    dPib += Pbb;
    // dPhr = -cp (T_min lr + dTs / 7 Pa^7 expm1(7 lr))
    double lrb = -cp * n.dTmr * dPhrb;
    Pa7b += -cp * dTs * (1.0 / 7.0) * E7 * dPhrb;
    lrb += -cp * dTs * (1.0 / 7.0) * Pa7 * 7.0 * (E7 + 1.0) * dPhrb;
    Pab += 7.0 * (Pa7 * iPa) * Pa7b;
    // up, d, lo
    double pgmb = 0.0, pgpb = 0.0, kab = 0.0, kbb = 0.0;
    pgpb += dtg * kb * 0.5 * gp * x10 * Ub;
    kbb += dtg * pgp * 0.5 * gp * x10 * Ub;
    x10b += dtg * pgp * kb * 0.5 * gp * Ub;
    const double c9 = 0.5 * g0 * x9;
    pgmb += dtg * ka * c9 * Db;
    kab += dtg * pgm * c9 * Db;
    pgpb += dtg * kb * c9 * Db;
    kbb += dtg * pgp * c9 * Db;
    x9b += dtg * (pgm * ka + pgp * kb) * 0.5 * g0 * Db;
    pgmb += dtg * ka * 0.5 * gm * x8 * Lb;
    kab += dtg * pgm * 0.5 * gm * x8 * Lb;
    x8b += dtg * pgm * ka * 0.5 * gm * Lb;
    double rab = -kap * kab, rbb = -kap * kbb;
    // e0, e1, r0, r1
    pgmb += dtg * kap * e0b;
    pgpb += dtg * kap * e1b;
    pgpb += dtg * dpb * r1b;
    const double dpbb = dtg * pgp * r1b;
    pgmb += dtg * dpa * r0b;
    const double dpab = dtg * pgm * r0b;
    const double buoyb = 0.5 * dtg * (r0b + r1b);
    // buoy = cp thbar dPi pgm
    const double thbarb = cp * dPi * pgm * buoyb;
    dPib += cp * thbar * pgm * buoyb;
    pgmb += cp * thbar * dPi * buoyb;
    // dPi = Pa expm1(lr), lr = kex log1p((pb - pa) / pa)
    Pab += Em * dPib;
    lrb += a.Pi * (Em + 1.0) * dPib;
    const double zb = n.dkex * lrb * ca_drcp(1.0 + z);
    double pbb = zb * ipa;
    double pab = -zb * b.p * (ipa * ipa);
    // thv = T / Pi
    Tab += 0.5 * thbarb * iPa;
    Pab -= 0.5 * thbarb * thva * iPa;
    Tbb += 0.5 * thbarb * iPib;
    double Pib_b = -0.5 * thbarb * thvb * iPib;
    // dp = kap ((T0 cp - Phi) - K)
    const double Kab = -kap * dpab, Kbb = -kap * dpbb;
    // pgp = -pgm, pgm = 1 / rf, rf = (rho_a + rho_b) / 2
    pgmb -= pgpb;
    const double rfb = -pgm * pgm * pgmb;
    rab += 0.5 * rfb;
    rbb += 0.5 * rfb;
    double oa[6] = {rab, 0.0, 0.0, 0.0, 0.0, 0.0}, obb[6] = {rbb, 0.0, 0.0, 0.0, 0.0, 0.0};
    cen_reverse(g, n, f - 1, a, x8, x9, Tab, pab, Pab, Kab, oa);
    cen_reverse(g, n, f, b, x9, x10, Tbb, pbb, Pib_b, Kbb, obb);
    pb[0] = oa[0];
    pb[1] = oa[1];
    pb[2] = oa[2];
    pb[3] = oa[3];
    pb[4] = obb[0];
    pb[5] = obb[1];
    pb[6] = obb[2];
    pb[7] = obb[3];
    pb[8] = x8b + oa[4];
    pb[9] = x9b + oa[5] + obb[4];
    pb[10] = x10b + obb[5];
    return v;
}

// --------------------------------------------------------------- LDS layout
template <int NF>
struct AdjSharedT {
    static constexpr int NFv = NF;
    ens::SharedT<NF> S;       // forward recompute
    float Ut[3][5][NF];       // U~_i of the step being reversed
    float Us[3][5][NF];       // U_i
    float ub[5][NF];          // cotangent of the step's start state u
    float feb[4][5][NF];      // of F_E_j
    float fib[3][3][NF];      // of F_I_i (rho, rho e_tot, u_3)
    float Ub[5][NF];          // of U_i, then of the pre-correction state V_i
    float Utb[5][NF];         // of U~_i
    float W[24][NF];          // scratch of the local reverse passes
    double lam[NF];           // transposed Schur solution
    double red[2][128];       // block reduction of the parameter cotangents (LPM <= 128)
};

// The part of AdjSharedT one implicit stage's reverse uses (newton_vjp alone)
template <int NF>
struct NewtonAdjShared {
    static constexpr int NFv = NF;
    ens::StageShared<NF> S;
    float Ut[1][5][NF];
    float Ub[5][NF];
    float Utb[5][NF];
    float W[24][NF];
    double lam[NF];
};

__device__ __forceinline__ void sync() { __syncthreads(); }

// Reverse of ens::explicit_tendency at Y: given the tendency cotangent Tb,
// adds the state cotangent to Yb and returns this thread's D0 cotangent.
template <int LPM, int NF>
__device__ double explicit_vjp(AdjSharedT<NF> &A, const DevGeom &g, const DevNum &n, const float (*Y)[NF],
                              const float (*Tb)[NF], float (*Yb)[NF], int tid) {
    const int nl = g.nlev;
    float (*W)[NF] = A.W;
    const float f3n = g.f3n[0], f12a = g.f12a[0], f12b = g.f12b[0];
    const float g11 = g.gu11[0], g12 = g.gu12[0], g22 = g.gu22[0];
    float Kf = 0.0f;
    bool kfree = false;
    if (n.vdiff) {
        const float D0 = g.D0_col[0];
        const float Kh = fmaxf(D0, 1.1920929e-07f);
        Kf = 1.0f / (0.5f * (1.0f / Kh + 1.0f / Kh));
        kfree = D0 > 1.1920929e-07f;
    }
    double kbar = 0.0;
    // centres: forward intermediates
    // W: 0 ct1, 1 ct2, 2 T (clamped flag in sign of 3), 3 unclamped, 4 sd, 5 rhoJ, 6 irJ, 7 G1 (faces), 8 G2,
    //    9 Fs, 10 Fu0, 11 Fu1 (faces); 12.. face -> centre contributions
    for (int k = tid; k < nl; k += LPM) {
        const float rho = Y[0][k], u1 = Y[1][k], u2 = Y[2][k], rhoe = Y[3][k];
        const float u3a = k == 0 ? 0.0f : Y[4][k], u3b = k + 1 >= nl ? 0.0f : Y[4][k + 1];
        const float is2 = g.is2c[k];
        const float ct1 = is2 * (g11 * u1 + g12 * u2), ct2 = is2 * (g12 * u1 + g22 * u2);
        const float Kh = 0.5f * (u1 * ct1 + u2 * ct2);
        const float Kv = 0.25f * (u3a * g.g33f[k] * u3a + u3b * g.g33f[k + 1] * u3b);
        const double e_tot = (double)rhoe * ca_drcp((double)rho);
        const double Tu = n.dT_0 + (e_tot - (double)(Kh + Kv) - g.Phic_d[k] + n.dR_d * n.dT_0) * n.dicv_d;
        const float Tf = (float)fmax(Tu, n.dTmin);
        const float J = g.Jh[0] * g.Jvc[k];
        W[0][k] = ct1;
        W[1][k] = ct2;
        W[3][k] = Tu > n.dTmin ? 1.0f : 0.0f;
        W[4][k] = n.cp_d * (Tf - n.T_0) + g.Phic[k];
        W[5][k] = rho * J;
        W[6][k] = ca_rcp(rho * J);
    }
    sync();
    // faces: forward face quantities
    for (int f = tid; f <= nl; f += LPM) {
        float G1 = 0.0f, G2 = 0.0f, Fs = 0.0f, Fu0 = 0.0f, Fu1 = 0.0f;
        if (f >= 1 && f <= nl - 1) {
            const int a = f - 1, b = f;
            const float Jf = g.Jh[0] * g.sf[f] * g.sf[f] * g.Jvf[f], iJf = g.iJh[0] * g.isJf[f];
            const float fu3 = Y[4][f] * g.g33f[f];
            const float A1 = f12a * g.isf[f] - (Y[2][b] - Y[2][a]) * iJf;
            const float A2 = f12b * g.isf[f] + (Y[1][b] - Y[1][a]) * iJf;
            const float b3 = (0.5f * (W[5][a] + W[5][b])) * fu3;
            G1 = Jf * (A2 * b3);
            G2 = Jf * (0.0f - A1 * b3);
            if (n.vdiff) {
                const float rK = (0.5f * (Y[0][a] + Y[0][b])) * Kf;
                const float iJv = g.iJvf[f];
                Fs = Jf * (iJv * iJv * (-(rK * (W[4][b] - W[4][a]))));
                const float fac = Jf * iJv * (-2.0f * rK) * 0.5f;
                Fu0 = fac * (Y[1][b] - Y[1][a]) * iJv;
                Fu1 = fac * (Y[2][b] - Y[2][a]) * iJv;
            }
        }
        W[7][f] = G1;
        W[8][f] = G2;
        W[9][f] = Fs;
        W[10][f] = Fu0;
        W[11][f] = Fu1;
    }
    sync();
    // faces: reverse of the face quantities; contributions to the centre
    // below (W[12..17]: rho, u1, u2, Kh, Kv, sd) and above (W[18..23]); the
    // face's own u_3 cotangent is added to Yb directly (one writer per face)
    for (int f = tid; f <= nl; f += LPM) {
        float ca[6] = {0, 0, 0, 0, 0, 0}, cb[6] = {0, 0, 0, 0, 0, 0};
        if (f >= 1 && f <= nl - 1) {
            const int a = f - 1, b = f;
            const float Jf = g.Jh[0] * g.sf[f] * g.sf[f] * g.Jvf[f], iJf = g.iJh[0] * g.isJf[f];
            const float fu3 = Y[4][f] * g.g33f[f];
            const float A1 = f12a * g.isf[f] - (Y[2][b] - Y[2][a]) * iJf;
            const float A2 = f12b * g.isf[f] + (Y[1][b] - Y[1][a]) * iJf;
            const float rJf = 0.5f * (W[5][a] + W[5][b]);
            const float b3 = rJf * fu3;
            const float i1 = 0.5f * (W[0][a] + W[0][b]), i2 = 0.5f * (W[1][a] + W[1][b]);
            // cotangents of G1, G2, Fs, Fu0, Fu1 from the centre completions of a and b
            const float iJca = g.iJh[0] * g.isJc[a], iJcb = g.iJh[0] * g.isJc[b];
            const float G1b = -0.5f * (W[6][a] * Tb[1][a] + W[6][b] * Tb[1][b]);
            const float G2b = -0.5f * (W[6][a] * Tb[2][a] + W[6][b] * Tb[2][b]);
            const float u3b_ = Tb[4][f];
            // du3 = -(Jf (A1 i2 - A2 i1) + (Kh_b - Kh_a) + (Kv_b - Kv_a))
            float A1b = -Jf * i2 * u3b_, A2b = Jf * i1 * u3b_;
            const float i2b = -Jf * A1 * u3b_, i1b = Jf * A2 * u3b_;
            ca[3] += u3b_;
            cb[3] -= u3b_;
            ca[4] += u3b_;
            cb[4] -= u3b_;
            // G1 = Jf A2 b3, G2 = -Jf A1 b3
            A2b += Jf * b3 * G1b;
            A1b += -Jf * b3 * G2b;
            const float b3b = Jf * A2 * G1b - Jf * A1 * G2b;
            const float rJfb = fu3 * b3b;
            Yb[4][f] += g.g33f[f] * rJf * b3b;
            // rJf = (rhoJ_a + rhoJ_b) / 2 -> rho
            ca[0] += 0.5f * rJfb * g.Jh[0] * g.Jvc[a];
            cb[0] += 0.5f * rJfb * g.Jh[0] * g.Jvc[b];
            // i1, i2 -> ct of a, b -> u via is2 G
            {
                const float c1 = 0.5f * i1b, c2 = 0.5f * i2b;
                const float isa = g.is2c[a], isb = g.is2c[b];
                ca[1] += isa * (g11 * c1 + g12 * c2);
                ca[2] += isa * (g12 * c1 + g22 * c2);
                cb[1] += isb * (g11 * c1 + g12 * c2);
                cb[2] += isb * (g12 * c1 + g22 * c2);
            }
            // A1 = f12a/s - (u2_b - u2_a)/Jf, A2 = f12b/s + (u1_b - u1_a)/Jf
            cb[2] += -iJf * A1b;
            ca[2] += iJf * A1b;
            cb[1] += iJf * A2b;
            ca[1] += -iJf * A2b;
            if (n.vdiff) {
                const float Fsb = -iJca * Tb[3][a] + iJcb * Tb[3][b];
                const float irA = ca_rcp(Y[0][a]), irB = ca_rcp(Y[0][b]);
                const float Fu0b = -iJca * irA * Tb[1][a] + iJcb * irB * Tb[1][b];
                const float Fu1b = -iJca * irA * Tb[2][a] + iJcb * irB * Tb[2][b];
                const float rK = (0.5f * (Y[0][a] + Y[0][b])) * Kf;
                const float iJv = g.iJvf[f];
                const float c = Jf * iJv * iJv;
                float rKb = -c * (W[4][b] - W[4][a]) * Fsb;
                cb[5] += -c * rK * Fsb;
                ca[5] += c * rK * Fsb;
                rKb += -c * (Y[1][b] - Y[1][a]) * Fu0b - c * (Y[2][b] - Y[2][a]) * Fu1b;
                cb[1] += -c * rK * Fu0b;
                ca[1] += c * rK * Fu0b;
                cb[2] += -c * rK * Fu1b;
                ca[2] += c * rK * Fu1b;
                ca[0] += 0.5f * Kf * rKb;
                cb[0] += 0.5f * Kf * rKb;
                if (kfree) kbar += (double)(0.5f * (Y[0][a] + Y[0][b]) * rKb);
            }
        }
#pragma unroll
        for (int q = 0; q < 6; ++q) {
            W[12 + q][f] = ca[q];
            W[18 + q][f] = cb[q];
        }
    }
    sync();
    // centres: gather (face k: this centre is above it; face k + 1: below),
    // local terms, then the chain through T, K, ct to the state; the Kv
    // cotangent goes to the u_3 of faces k, k + 1 through W[0], W[1]
    for (int k = tid; k < nl; k += LPM) {
        const float rho = Y[0][k], u1 = Y[1][k], u2 = Y[2][k], rhoe = Y[3][k];
        const float u3a = k == 0 ? 0.0f : Y[4][k], u3b = k + 1 >= nl ? 0.0f : Y[4][k + 1];
        float rhob = W[18][k] + W[12][k + 1];
        float u1b = W[19][k] + W[13][k + 1];
        float u2b = W[20][k] + W[14][k + 1];
        float Khb = W[21][k] + W[15][k + 1];
        float Kvb = W[22][k] + W[16][k + 1];
        const float sdb = W[23][k] + W[17][k + 1];
        const float J = g.Jh[0] * g.Jvc[k];
        const float irJ = W[6][k];
        // T1 = J a3 ct2 - b u1 - (G1_k + G1_k+1)/2 irJ - e0 irho (and T2 likewise)
        const float irJb = -0.5f * (W[7][k] + W[7][k + 1]) * Tb[1][k] - 0.5f * (W[8][k] + W[8][k + 1]) * Tb[2][k];
        rhob += -irJb * irJ * irJ * J;
        if (n.vdiff) {
            const float iJc = g.iJh[0] * g.isJc[k];
            const float e0 = (W[10][k + 1] - W[10][k]) * iJc, e1 = (W[11][k + 1] - W[11][k]) * iJc;
            const float irho = ca_rcp(rho);
            const float irhob = -e0 * Tb[1][k] - e1 * Tb[2][k];
            rhob += -irhob * irho * irho;
        }
        const float a3 = f3n * g.iJvc[k];
        float ct1b = -J * a3 * Tb[2][k], ct2b = J * a3 * Tb[1][k];
        if (n.rayleigh && n.alpha_uh != 0.0f) {
            const float beta = n.alpha_uh * g.zeta_c[k];
            u1b += -beta * Tb[1][k];
            u2b += -beta * Tb[2][k];
        }
        // sd = cp (T - T_0) + Phi;  T = T_0 + (rhoe/rho - K - Phi + R T_0)/cv (unclamped)
        const float Tb_ = n.cp_d * sdb;
        float rhoeb = 0.0f;
        if (W[3][k] > 0.0f) {
            const float icv = (float)n.dicv_d;
            rhoeb += Tb_ * icv / rho;
            rhob += -Tb_ * icv * rhoe / (rho * rho);
            Khb += -Tb_ * icv;
            Kvb += -Tb_ * icv;
        }
        // Kh = u . ct / 2 (ct = is2 G u): d Kh / d u = ct
        u1b += Khb * W[0][k];
        u2b += Khb * W[1][k];
        {
            const float is2 = g.is2c[k];
            u1b += is2 * (g11 * ct1b + g12 * ct2b);
            u2b += is2 * (g12 * ct1b + g22 * ct2b);
        }
        Yb[0][k] += rhob;
        Yb[1][k] += u1b;
        Yb[2][k] += u2b;
        Yb[3][k] += rhoeb;
        // Kv = (u3a^2 g_k + u3b^2 g_k+1) / 4
        W[0][k] = 0.5f * g.g33f[k] * u3a * Kvb;      // to face k
        W[1][k] = 0.5f * g.g33f[k + 1] * u3b * Kvb;  // to face k + 1
    }
    sync();
    for (int f = tid; f <= nl; f += LPM) {
        if (f >= 1 && f <= nl - 1) Yb[4][f] += W[0][f] + W[1][f - 1];
    }
    sync();
    return kbar;
}

// derivative of ca_vl_slope(a0, a1, a2) on the branch the forward takes
__device__ __forceinline__ void vl_slope_grad(float a0, float a1, float a2, float &g0, float &g1, float &g2) {
    g0 = g1 = g2 = 0.0f;
    const float d = 0.5f * ((a1 - a0) + (a2 - a1));
    if (!(d > 0.0f) && !(d < 0.0f)) return;
    const float sg = d > 0.0f ? 1.0f : -1.0f;
    const float mn = fminf(fminf(a0, a1), a2), mx = fmaxf(fmaxf(a0, a1), a2);
    const float l1 = fabsf(d), l2 = 2.0f * (a1 - mn), l3 = 2.0f * (mx - a1);
    const float inner = fminf(l2, l3);
    if (l1 <= inner) {  // lim = |d|, s = d
        g0 = -0.5f;
        g2 = 0.5f;
        return;
    }
    if (l2 <= l3) {  // s = sg * 2 (a1 - mn)
        g1 += 2.0f * sg;
        if (mn == a0) g0 -= 2.0f * sg;
        else if (mn == a1) g1 -= 2.0f * sg;
        else g2 -= 2.0f * sg;
    } else {  // s = sg * 2 (mx - a1)
        g1 -= 2.0f * sg;
        if (mx == a0) g0 += 2.0f * sg;
        else if (mx == a1) g1 += 2.0f * sg;
        else g2 += 2.0f * sg;
    }
}

// d upwind_product / d (v, c2, c3, c4, c5)
__device__ __forceinline__ void upwind_grad(int scheme, int f, int nl, float v, float dt, float c2, float c3, float c4,
                                            float c5, float &gv, float *gc) {
    gc[0] = gc[1] = gc[2] = gc[3] = 0.0f;
    if (scheme == 0) {
        gv = 0.5f * (c3 + c4);
        gc[1] = gc[2] = 0.5f * v;
        return;
    }
    if (scheme == 1 || f == 1 || f == nl - 1) {
        if (v >= 0.0f) {
            gv = c3;
            gc[1] = v;
        } else {
            gv = c4;
            gc[2] = v;
        }
        return;
    }
    float s0, s1, s2;
    if (v >= 0.0f) {
        const float s = ca_vl_slope(c2, c3, c4);
        vl_slope_grad(c2, c3, c4, s0, s1, s2);
        const float w = (1.0f - v * dt) * 0.5f;
        gv = c3 + w * s + v * (-dt * s * 0.5f);
        gc[0] = v * w * s0;
        gc[1] = v * (1.0f + w * s1);
        gc[2] = v * w * s2;
    } else {
        const float s = ca_vl_slope(c3, c4, c5);
        vl_slope_grad(c3, c4, c5, s0, s1, s2);
        const float w = (1.0f + v * dt) * 0.5f;
        gv = c4 - w * s + v * (-dt * s * 0.5f);
        gc[1] = -v * w * s0;
        gc[2] = v * (1.0f - w * s1);
        gc[3] = -v * w * s2;
    }
}

// dh/d(rho, u1, u2, rhoe, u3a, u3b) of a centre (cv_make's h = e_tot + R_d T)
__device__ __forceinline__ void h_grad(const DevGeom &g, const DevNum &n, int k, float rho, float u1, float u2,
                                       float rhoe, float u3a, float u3b, float hb, float *out) {
    const float is2 = g.is2c[k];
    const float ct1 = is2 * (g.gu11[0] * u1 + g.gu12[0] * u2), ct2 = is2 * (g.gu12[0] * u1 + g.gu22[0] * u2);
    const float Kh = 0.5f * (u1 * ct1 + u2 * ct2);
    const float Kv = 0.25f * (u3a * g.g33f[k] * u3a + u3b * g.g33f[k + 1] * u3b);
    const double e_tot = (double)rhoe / (double)rho;
    const double Tu = n.dT_0 + (e_tot - (double)(Kh + Kv) - g.Phic_d[k] + n.dR_d * n.dT_0) * n.dicv_d;
    const float Rcv = (float)(n.dR_d * n.dicv_d);
    const float de = Tu > n.dTmin ? 1.0f + Rcv : 1.0f;  // dh/de_tot
    const float dK = Tu > n.dTmin ? -Rcv : 0.0f;         // dh/dK
    out[0] += hb * de * (-rhoe / (rho * rho));
    out[3] += hb * de / rho;
    const float Kb = hb * dK;
    out[1] += Kb * ct1;
    out[2] += Kb * ct2;
    out[4] += Kb * 0.5f * g.g33f[k] * u3a;
    out[5] += Kb * 0.5f * g.g33f[k + 1] * u3b;
}

// The internals of ens::newton_stage that its reverse reads, rebuilt from
// the stage's recorded result instead of by re-solving the Schur system.
// On entry S.U holds U~, and rhoN, u3N (LDS) rho and u_3 of the recorded
// result U.  The face increment is x = u~_3 - u_3: what the stage subtracted,
// up to the rounding of that float subtraction (exact when the two are
// within a factor 2).  rho e_tot before the upwind correction is
// (rho e_tot)~ - de(x) with de = e0 x_k + e1 x_k+1 - Re, as in the stage.
// On exit S.U holds U with rho e_tot before the correction, and S.nw.c, cd,
// row, x, ve (and hr, rJ with upwinding) hold what newton_stage leaves.
template <int LPM, class SH>
__device__ void newton_replay(SH &S, const DevGeom &g, const DevNum &n, float dtg, const float *rhoN,
                              const float *u3N, int tid) {
    const int nl = g.nlev;
    ens::stage_thermo<LPM>(S, g, n, tid);
    for (int f = tid; f <= nl; f += LPM) S.nw.x[f] = (f >= 1 && f <= nl - 1) ? S.U[4][f] - u3N[f] : 0.0f;
    sync();
    ens::stage_rows<LPM>(S, g, dtg, tid);
    sync();
    for (int k = tid; k < nl; k += LPM) {
        const float de = S.nw.row[2][k] * S.nw.x[k] + S.nw.row[3][k] * S.nw.x[k + 1] - S.nw.row[5][k];
        S.U[0][k] = rhoN[k];
        S.U[3][k] -= de;
        S.nw.ve[k] = S.U[3][k];
    }
    for (int f = tid; f <= nl; f += LPM) S.U[4][f] = (f >= 1 && f <= nl - 1) ? u3N[f] : 0.0f;
    sync();
    if (n.upw != 0) {
        ens::stage_enthalpy<LPM>(S, g, n, tid);
        sync();
    }
}

// Reverse of one implicit stage (ens::newton_stage) given its recorded
// result (rhoN, u3N: see newton_replay).  On entry S.U holds U~ (= Ut) and
// A.Ub the cotangent of the result U; on exit A.Utb holds that of U~.
// Returns this thread's share of the alpha_w cotangent.
template <int LPM, class AS>
__device__ double newton_vjp(AS &A, const DevGeom &g, const DevNum &n, float dtg, const float (*Ut)[AS::NFv],
                            const float *rhoN, const float *u3N, int tid) {
    constexpr int NF = AS::NFv;
    auto &S = A.S;
    const int nl = g.nlev;
    float (*W)[NF] = A.W;
    newton_replay<LPM>(S, g, n, dtg, rhoN, u3N, tid);  // S.U = U_i, S.nw = internals; V_i (rho e) in S.nw.ve
    double awbar = 0.0;
    // ---- upwind correction: U_e = V_e + dtg corr(V); V = U except rho e
    if (n.upw != 0) {
        // centres: V cotangent starts as U cotangent (corr reads V, not U)
        // faces: cotangent of dF_f from corr_k = -(dF_k+1 - dF_k) iJ_k
        for (int f = tid; f <= nl; f += LPM) {
            float dFb = 0.0f;
            if (f >= 1 && f <= nl - 1) {
                dFb = dtg * (-A.Ub[3][f - 1] * S.nw.c[ens::NC_IJ][f - 1] + A.Ub[3][f] * S.nw.c[ens::NC_IJ][f]);
            }
            W[0][f] = dFb;
        }
        sync();
        // faces: dF = rJf (upw - cen), rJf = (rJ_f-1 + rJ_f)/2, v = u3 g33
        // contributions: W[1..4] h of centres f-2..f+1, W[5] rJ below, W[6] rJ above; u3 direct
        for (int f = tid; f <= nl; f += LPM) {
            float hc[4] = {0, 0, 0, 0}, rJa = 0.0f, rJbv = 0.0f;
            if (f >= 1 && f <= nl - 1) {
                const float dFb = W[0][f];
                const float href = S.nw.hr[0];
                const float v = S.U[4][f] * g.g33f[f];
                const float rJf = 0.5f * (S.nw.rJ[f - 1] + S.nw.rJ[f]);
                const float h2 = f >= 2 ? S.nw.hr[f - 2] - href : 0.0f;
                const float h1 = S.nw.hr[f - 1] - href, h3 = S.nw.hr[f] - href;
                const float h4 = f + 1 <= nl - 1 ? S.nw.hr[f + 1] - href : 0.0f;
                const float upw = upwind_product(n.upw, f, nl, v, n.dt, h2, h1, h3, h4);
                const float cen = upwind_product(0, f, nl, v, n.dt, 0.0f, h1, h3, 0.0f);
                float gvu, gcu[4], gvc, gcc[4];
                upwind_grad(n.upw, f, nl, v, n.dt, h2, h1, h3, h4, gvu, gcu);
                upwind_grad(0, f, nl, v, n.dt, 0.0f, h1, h3, 0.0f, gvc, gcc);
                const float rJfb = dFb * (upw - cen);
                rJa = 0.5f * rJfb;
                rJbv = 0.5f * rJfb;
                const float vb = dFb * rJf * (gvu - gvc);
                A.Ub[4][f] += vb * g.g33f[f];
#pragma unroll
                for (int q = 0; q < 4; ++q) hc[q] = dFb * rJf * (gcu[q] - gcc[q]);
                if (f < 2) hc[0] = 0.0f;
                if (f + 1 > nl - 1) hc[3] = 0.0f;
            }
            W[1][f] = hc[0];
            W[2][f] = hc[1];
            W[3][f] = hc[2];
            W[4][f] = hc[3];
            W[5][f] = rJa;
            W[6][f] = rJbv;
        }
        sync();
        // centres: gather h and rJ cotangents of centre k: from faces k+2 (as
        // f-2), k+1 (as f-1), k (as f), k-1 (as f+1); rJ from faces k+1, k
        for (int k = tid; k < nl; k += LPM) {
            float hb = W[2][k + 1] + W[3][k];
            if (k + 2 <= nl) hb += W[1][k + 2];
            if (k >= 1) hb += W[4][k - 1];
            const float rJb = W[5][k + 1] + W[6][k];
            float out[6] = {0, 0, 0, 0, 0, 0};
            const float Ve = S.nw.ve[k];
            const float u3a = k == 0 ? 0.0f : S.U[4][k], u3b = k + 1 >= nl ? 0.0f : S.U[4][k + 1];
            h_grad(g, n, k, S.U[0][k], S.U[1][k], S.U[2][k], Ve, u3a, u3b, hb, out);
            out[0] += rJb * S.nw.c[ens::NC_J][k];
            A.Ub[0][k] += out[0];
            A.Ub[1][k] += out[1];
            A.Ub[2][k] += out[2];
            A.Ub[3][k] += out[3];
            W[7][k] = out[4];  // to face k
            W[8][k] = out[5];  // to face k + 1
        }
        sync();
        for (int f = tid; f <= nl; f += LPM)
            if (f >= 1 && f <= nl - 1) A.Ub[4][f] += W[7][f] + W[8][f - 1];
        sync();
    }
    // ---- Newton update: V = U~ - delta; A.Ub now holds the V cotangent.
    // delta: dr_k = a0 x_k + a1 x_k+1 - Rr, de_k = e0 x_k + e1 x_k+1 - Re,
    // x = S^-1 q on faces 0 .. nl - 1 (x_nl = 0)
    // identity path, and the Schur rows (L, D, U) with x_bar in sys[1]
    for (int f = tid; f <= nl; f += LPM) {
        for (int p = 0; p < 4; ++p)
            if (f < nl) A.Utb[p][f] = A.Ub[p][f];
        A.Utb[4][f] = (f >= 1 && f <= nl - 1) ? A.Ub[4][f] : 0.0f;
    }
    for (int f = tid; f < nl; f += LPM) {
        double L = 0.0, D, U = 0.0;
        float xb = 0.0f;
        if (f >= 1) {
            const CV a = ens::cv_from(S, f - 1), c = ens::cv_from(S, f);
            const float u3m = f - 1 >= 1 ? Ut[4][f - 1] : 0.0f;
            const float u3 = Ut[4][f];
            const float u3p = f + 1 <= nl - 1 ? Ut[4][f + 1] : 0.0f;
            const FaceRow fr = jac_face_row(g, n, f, 0, a, c, u3m, u3, u3p, dtg);
            const float pa0 = S.nw.row[0][f - 1], pa1 = S.nw.row[1][f - 1];
            const float pe0 = S.nw.row[2][f - 1], pe1 = S.nw.row[3][f - 1];
            const float a0 = S.nw.row[0][f], e0 = S.nw.row[2][f], a1 = S.nw.row[1][f], e1 = S.nw.row[3][f];
            L = fr.lo + fr.r0 * pa0 + fr.e0 * pe0;
            D = fr.d + fr.r0 * pa1 + fr.e0 * pe1 + fr.r1 * a0 + fr.e1 * e0;
            U = f + 1 <= nl - 1 ? fr.up + fr.r1 * a1 + fr.e1 * e1 : 0.0;
            xb = f <= nl - 1 ? -A.Ub[4][f] : 0.0f;
            // centre f - 1: dr_{f-1} has a1 x_f, de_{f-1} has e1 x_f
            xb += (-A.Ub[0][f - 1]) * pa1 + (-A.Ub[3][f - 1]) * pe1;
        } else {
            D = -dtg * beta_w(g, n, 0, 0) - 1.0f;
        }
        // centre f: dr_f has a0 x_f, de_f has e0 x_f
        xb += (-A.Ub[0][f]) * S.nw.row[0][f] + (-A.Ub[3][f]) * S.nw.row[2][f];
        S.nw.sys[1][0][f] = L;
        S.nw.sys[1][1][f] = D;
        S.nw.sys[1][2][f] = U;
        S.nw.sys[1][3][f] = xb;
    }
    sync();
    // transposed rows: L'_f = U_{f-1}, D'_f = D_f, U'_f = L_{f+1}
    for (int f = tid; f < nl; f += LPM) {
        S.nw.sys[0][0][f] = f >= 1 ? S.nw.sys[1][2][f - 1] : 0.0;
        S.nw.sys[0][1][f] = S.nw.sys[1][1][f];
        S.nw.sys[0][2][f] = f + 1 < nl ? S.nw.sys[1][0][f + 1] : 0.0;
        S.nw.sys[0][3][f] = S.nw.sys[1][3][f];
    }
    sync();
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
    for (int f = tid; f <= nl; f += LPM) A.lam[f] = f < nl ? S.nw.sys[src][3][f] * ca_drcp(S.nw.sys[src][1][f]) : 0.0;
    sync();
    // faces 1 .. nl - 1: cotangents of the Schur row entries and rhs, then
    // of the 11 values the face row reads (face_vjp); W[0..3] centre
    // below, W[4..7] centre above, W[8..10] u_3 of faces f-1, f, f+1;
    // W[11..14] the row's r0, e0, r1, e1 for the centre-row pass
    for (int f = tid; f <= nl; f += LPM) {
        float pb[NP];
#pragma unroll
        for (int i = 0; i < NP; ++i) pb[i] = 0.0f;
        float r0v = 0.0f, e0v = 0.0f, r1v = 0.0f, e1v = 0.0f;
        if (f >= 1 && f <= nl - 1) {
            float prim[NP] = {Ut[0][f - 1], Ut[1][f - 1], Ut[2][f - 1], Ut[3][f - 1],
                              Ut[0][f],     Ut[1][f],     Ut[2][f],     Ut[3][f],
                              f - 1 >= 1 ? Ut[4][f - 1] : 0.0f, Ut[4][f], f + 1 <= nl - 1 ? Ut[4][f + 1] : 0.0f};
            const double lam = A.lam[f];
            const double xm = S.nw.x[f - 1], x0 = S.nw.x[f], xp = S.nw.x[f + 1];
            const bool hasU = f + 1 <= nl - 1;
            const double Lb = -lam * xm, Db = -lam * x0, Ub = hasU ? -lam * xp : 0.0, qb = lam;
            const double pa0 = S.nw.row[0][f - 1], pa1 = S.nw.row[1][f - 1], pe0 = S.nw.row[2][f - 1],
                         pe1 = S.nw.row[3][f - 1], pRr = S.nw.row[4][f - 1], pRe = S.nw.row[5][f - 1];
            const double a0 = S.nw.row[0][f], a1 = S.nw.row[1][f], e0 = S.nw.row[2][f], e1 = S.nw.row[3][f],
                         Rr = S.nw.row[4][f], Re = S.nw.row[5][f];
            const double ob[8] = {Lb,
                                  Db,
                                  Ub,
                                  Lb * pa0 + Db * pa1 + qb * pRr,
                                  Db * a0 + Ub * a1 + qb * Rr,
                                  Lb * pe0 + Db * pe1 + qb * pRe,
                                  Db * e0 + Ub * e1 + qb * Re,
                                  qb};
            double pbd[NP];
            const double tpa[3] = {S.nw.cd[0][f - 1], S.nw.cd[1][f - 1], S.nw.cd[2][f - 1]};
            const double tpb[3] = {S.nw.cd[0][f], S.nw.cd[1][f], S.nw.cd[2][f]};
            const FaceVals o = face_vjp(g, n, f, dtg, prim, tpa, tpb, ob, pbd);
#pragma unroll
            for (int i = 0; i < NP; ++i) pb[i] = (float)pbd[i];
            if (f - 1 < 1) pb[8] = 0.0f;
            if (!hasU) pb[10] = 0.0f;
            if (n.rayleigh) awbar += g.zeta_f[f] * (-(double)dtg) * (Db + qb * (double)prim[9]);
            r0v = (float)o.r0;
            e0v = (float)o.e0;
            r1v = (float)o.r1;
            e1v = (float)o.e1;
        }
#pragma unroll
        for (int i = 0; i < NP; ++i) W[i][f] = pb[i];
        W[11][f] = r0v;
        W[12][f] = e0v;
        W[13][f] = r1v;
        W[14][f] = e1v;
    }
    sync();
    // centres: cotangents of the centre rows (a0, a1, e0, e1, Rr, Re) from
    // the Schur rows of faces k and k+1 and from dr, de; then the rows'
    // reverse onto rJ and h of centres k-1, k, k+1 (W[15..20]) and the u_3 of
    // faces k, k+1 (W[21], W[22])
    for (int k = tid; k < nl; k += LPM) {
        const double lk = A.lam[k], lk1 = k + 1 < nl ? A.lam[k + 1] : 0.0;
        const double xk = S.nw.x[k], xk1 = S.nw.x[k + 1];
        double a0b = 0.0, a1b = 0.0, e0b = 0.0, e1b = 0.0, Rrb = 0.0, Reb = 0.0;
        if (k >= 1) {  // face k: r1, e1 rows multiply centre k's a0, e0 (D) and a1, e1 (U)
            const double Dbk = -lk * xk, Ubk = (k + 1 <= nl - 1) ? -lk * xk1 : 0.0;
            const double r1 = W[13][k], e1f = W[14][k];
            a0b += Dbk * r1;
            e0b += Dbk * e1f;
            a1b += Ubk * r1;
            e1b += Ubk * e1f;
            Rrb += lk * r1;
            Reb += lk * e1f;
        }
        if (k + 1 <= nl - 1) {  // face k + 1: r0, e0 rows multiply centre k's a0, e0 (L) and a1, e1 (D)
            const double Lb1 = -lk1 * xk, Db1 = -lk1 * xk1;
            const double r0 = W[11][k + 1], e0f = W[12][k + 1];
            a0b += Lb1 * r0;
            e0b += Lb1 * e0f;
            a1b += Db1 * r0;
            e1b += Db1 * e0f;
            Rrb += lk1 * r0;
            Reb += lk1 * e0f;
        }
        const double drb = -A.Ub[0][k], deb = -A.Ub[3][k];
        a0b += drb * xk;
        a1b += drb * xk1;
        Rrb += -drb;
        e0b += deb * xk;
        e1b += deb * xk1;
        Reb += -deb;
        // the centre rows (k) as computed in ens::newton_stage
        const float iJ = S.nw.c[ens::NC_IJ][k];
        const float rJ = S.nw.c[ens::NC_RJ][k], h = S.nw.c[ens::NC_H][k];
        const double Fpb = -(double)dtg * iJ * Rrb, Fmb0 = (double)dtg * iJ * Rrb;
        const double Hpb = -(double)dtg * iJ * Reb, Hmb = (double)dtg * iJ * Reb;
        float rJlo = 0.0f, hlo = 0.0f, rJown = 0.0f, hown = 0.0f, rJup = 0.0f, hup = 0.0f, u3k = 0.0f, u3k1 = 0.0f;
        if (k >= 1) {
            const float rJfm = 0.5f * (S.nw.c[ens::NC_RJ][k - 1] + rJ);
            const float hfm = 0.5f * (S.nw.c[ens::NC_H][k - 1] + h);
            const float g3 = g.g33f[k];
            const float a0 = S.nw.row[0][k];
            const float Fm = rJfm * (Ut[4][k] * g3);
            const double a0t = a0b + e0b * hfm;
            const double hfmb = e0b * a0 + Hmb * Fm;
            const double Fmb = Fmb0 + Hmb * hfm;
            const double rJfmb = a0t * dtg * iJ * g3 + Fmb * Ut[4][k] * g3;
            u3k += (float)(Fmb * rJfm * g3);
            rJlo += (float)(0.5 * rJfmb);
            rJown += (float)(0.5 * rJfmb);
            hlo += (float)(0.5 * hfmb);
            hown += (float)(0.5 * hfmb);
        }
        if (k + 1 <= nl - 1) {
            const float rJfp = 0.5f * (rJ + S.nw.c[ens::NC_RJ][k + 1]);
            const float hfp = 0.5f * (h + S.nw.c[ens::NC_H][k + 1]);
            const float g3 = g.g33f[k + 1];
            const float a1 = S.nw.row[1][k];
            const float Fp = rJfp * (Ut[4][k + 1] * g3);
            const double a1t = a1b + e1b * hfp;
            const double hfpb = e1b * a1 + Hpb * Fp;
            const double Fpt = Fpb + Hpb * hfp;
            const double rJfpb = a1t * (-(double)dtg) * iJ * g3 + Fpt * Ut[4][k + 1] * g3;
            u3k1 += (float)(Fpt * rJfp * g3);
            rJown += (float)(0.5 * rJfpb);
            rJup += (float)(0.5 * rJfpb);
            hown += (float)(0.5 * hfpb);
            hup += (float)(0.5 * hfpb);
        }
        W[15][k] = rJlo;
        W[16][k] = hlo;
        W[17][k] = rJown;
        W[18][k] = hown;
        W[19][k] = rJup;
        W[20][k] = hup;
        W[21][k] = u3k;
        W[22][k] = u3k1;
    }
    sync();
    // centres: gather rJ, h cotangents; chain through rJ = rho J and h(U~);
    // add the face rows' centre cotangents (face k: above, face k+1: below)
    for (int k = tid; k < nl; k += LPM) {
        float rJb = W[17][k], hb = W[18][k];
        if (k >= 1) {
            rJb += W[19][k - 1];
            hb += W[20][k - 1];
        }
        if (k + 1 < nl) {
            rJb += W[15][k + 1];
            hb += W[16][k + 1];
        }
        float out[6] = {0, 0, 0, 0, 0, 0};
        const float u3a = k == 0 ? 0.0f : Ut[4][k], u3b = k + 1 >= nl ? 0.0f : Ut[4][k + 1];
        h_grad(g, n, k, Ut[0][k], Ut[1][k], Ut[2][k], Ut[3][k], u3a, u3b, hb, out);
        out[0] += rJb * S.nw.c[ens::NC_J][k];
        for (int p = 0; p < 4; ++p) {
            float v = out[p];
            if (k >= 1) v += W[4 + p][k];
            if (k + 1 <= nl - 1) v += W[p][k + 1];
            A.Utb[p][k] += v;
        }
        W[21][k] += out[4];
        W[22][k] += out[5];
    }
    sync();
    for (int f = tid; f <= nl; f += LPM) {
        if (f >= 1 && f <= nl - 1) {
            float v = W[21][f] + W[22][f - 1] + W[9][f];
            if (f + 1 <= nl - 1) v += W[8][f + 1];
            if (f - 1 >= 1) v += W[10][f - 1];
            A.Utb[4][f] += v;
        }
    }
    sync();
    return awbar;
}

// Reverse of one step.  On entry A.S.u holds the step's start state u, A.Ut
// and A.Us the stage states of that step (step_once with recording), and
// A.ub the cotangent of the step's result; on exit A.ub is the cotangent of u.
// kb, ab accumulate this thread's D0 and alpha_w cotangents.
template <int LPM, int NF>
__device__ void reverse_step(AdjSharedT<NF> &A, const DevGeom &g, const DevNum &n, const EnsTab &tab, float dt, int tid,
                             double &kb, double &ab) {
    const int nl = g.nlev;
    auto centre = [&](int p, int k) { return p < 4 ? k < nl : (k >= 1 && k <= nl - 1); };
    const int ci3[5] = {0, -1, -1, 1, 2};
    // u' = u + dt sum_j b_j (F_E_j + F_I_j); u'_3 = 0 on the boundary faces
    for (int q = tid; q < 5 * NF; q += LPM) {
        const int p = q / NF, k = q - p * NF;
        if (p == 4 && (k == 0 || k == nl)) A.ub[4][k] = 0.0f;
        const float v = centre(p, k) ? A.ub[p][k] : 0.0f;
        for (int j = 0; j < 4; ++j) A.feb[j][p][k] = dt * tab.b[j] * v;
        if (ci3[p] >= 0)
            for (int j = 1; j < 4; ++j) A.fib[j - 1][ci3[p]][k] = dt * tab.b[j] * v;
    }
    sync();
    for (int i = 3; i >= 1; --i) {
        const float dtg = dt * tab.aI[i][i];
        const float idtg = 1.0f / dtg;
        for (int q = tid; q < 5 * NF; q += LPM) A.Ub[q / NF][q % NF] = 0.0f;
        sync();
        // F_E_i = T_E(U_i)
        kb += explicit_vjp<LPM>(A, g, n, A.Us[i - 1], A.feb[i], A.Ub, tid);
        // F_I_i = (U_i - U~_i) / dtg
        for (int q = tid; q < 5 * NF; q += LPM) {
            const int p = q / NF, k = q - p * NF;
            if (ci3[p] >= 0 && centre(p, k)) A.Ub[p][k] += A.fib[i - 1][ci3[p]][k] * idtg;
        }
        // U_i = Newton(U~_i): rebuild from U~_i
        for (int q = tid; q < 5 * NF; q += LPM) {
            const int p = q / NF, k = q - p * NF;
            A.S.U[p][k] = A.Ut[i - 1][p][k];
        }
        sync();
        ab += newton_vjp<LPM>(A, g, n, dtg, A.Ut[i - 1], A.Us[i - 1][0], A.Us[i - 1][4], tid);
        // U~_i = u + dt sum_j<i (aE F_E_j + aI F_I_j)
        for (int q = tid; q < 5 * NF; q += LPM) {
            const int p = q / NF, k = q - p * NF;
            if (!centre(p, k)) continue;
            float tb = A.Utb[p][k];
            if (ci3[p] >= 0) tb -= A.fib[i - 1][ci3[p]][k] * idtg;
            A.ub[p][k] += tb;
            for (int j = 0; j < i; ++j) {
                A.feb[j][p][k] += dt * tab.aE[i][j] * tb;
                if (j >= 1 && ci3[p] >= 0) A.fib[j - 1][ci3[p]][k] += dt * tab.aI[i][j] * tb;
            }
        }
        sync();
    }
    // F_E_0 = T_E(u)
    kb += explicit_vjp<LPM>(A, g, n, A.S.u, A.feb[0], A.ub, tid);
    for (int f = tid; f <= nl; f += LPM)
        if (f == 0 || f == nl) A.ub[4][f] = 0.0f;
    sync();
}

}  // namespace adj

// Adjoint of nsteps ensemble steps: workgroup b reverses member order[b].
// Y0: initial states; YbN: cotangents of the states after nsteps steps.
// Outputs: Yb0 (cotangent of Y0), D0b and awb per member.  ckpt / seg:
// scratch of (ceil(nsteps / every) + 1) and `every` ensemble states; member
// m's part starts at that multiple of its state offset off[m] (the offsets
// are the prefix sums of the member state sizes, so the parts are disjoint).
template <int LPM, int NF>
__global__ void __launch_bounds__(LPM) k_ensemble_adjoint(const DevGeom *__restrict__ mg, DevNum n,
                                                          const int *__restrict__ order, const int64_t *__restrict__ off,
                                                          const float *__restrict__ Y0, const float *__restrict__ YbN,
                                                          float *__restrict__ Yb0, float *__restrict__ D0b,
                                                          float *__restrict__ awb, float *__restrict__ ckpt,
                                                          float *__restrict__ seg, EnsTab tab, float dt, int nsteps,
                                                          int every) {
    using namespace adj;
    static_assert(LPM <= 128, "the adjoint runs at most 128 lanes per member");
    __shared__ AdjSharedT<NF> A;
    auto &S = A.S;
    const int m = order[blockIdx.x];
    const DevGeom &g = mg[m];
    const int nl = g.nlev;
    const int tid = threadIdx.x;
    const int64_t Sm = 5 * (int64_t)nl + 1;
    const int nseg = (nsteps + every - 1) / every;
    float *ck = ckpt + (int64_t)(nseg + 1) * off[m];
    float *sg = seg + (int64_t)every * off[m];
    auto load = [&](const float *src) {  // member state -> S.u (u_3 filtered)
        for (int k = tid; k < nl; k += LPM)
            for (int p = 0; p < 4; ++p) S.u[p][k] = src[p * nl + k];
        for (int f = tid; f <= nl; f += LPM) S.u[4][f] = (f == 0 || f == nl) ? 0.0f : src[4 * nl + f];
        sync();
    };
    auto store = [&](float *dst) {
        for (int k = tid; k < nl; k += LPM)
            for (int p = 0; p < 4; ++p) dst[p * nl + k] = S.u[p][k];
        for (int f = tid; f <= nl; f += LPM) dst[4 * nl + f] = S.u[4][f];
    };
    // forward, storing a checkpoint every `every` steps
    load(Y0 + off[m]);
    store(ck);
    for (int st = 0; st < nsteps; ++st) {
        ens::step_once<LPM>(S, g, n, tab, dt, tid);
        if ((st + 1) % every == 0 && st + 1 < nsteps) store(ck + ((st + 1) / every) * Sm);
    }
    // seed
    const float *ybn = YbN + off[m];
    for (int k = tid; k < nl; k += LPM)
        for (int p = 0; p < 4; ++p) A.ub[p][k] = ybn[p * nl + k];
    for (int f = tid; f <= nl; f += LPM) A.ub[4][f] = (f == 0 || f == nl) ? 0.0f : ybn[4 * nl + f];
    sync();
    double kb = 0.0, ab = 0.0;
    for (int sgi = nseg - 1; sgi >= 0; --sgi) {
        const int len = min(every, nsteps - sgi * every);
        load(ck + sgi * Sm);
        for (int j = 0; j < len; ++j) {
            store(sg + j * Sm);
            if (j + 1 < len) ens::step_once<LPM>(S, g, n, tab, dt, tid);
        }
        sync();
        for (int j = len - 1; j >= 0; --j) {
            load(sg + j * Sm);
            ens::step_once<LPM>(S, g, n, tab, dt, tid, A.Ut, A.Us);
            load(sg + j * Sm);
            reverse_step<LPM>(A, g, n, tab, dt, tid, kb, ab);
        }
    }
    float *yb0 = Yb0 + off[m];
    for (int k = tid; k < nl; k += LPM)
        for (int p = 0; p < 4; ++p) yb0[p * nl + k] = A.ub[p][k];
    for (int f = tid; f <= nl; f += LPM) yb0[4 * nl + f] = A.ub[4][f];
    // parameter cotangents: block sum (fixed order: deterministic)
    A.red[0][tid] = kb;
    A.red[1][tid] = ab;
    sync();
    if (tid == 0) {
        double s0 = 0.0, s1 = 0.0;
        for (int t = 0; t < LPM; ++t) {
            s0 += A.red[0][t];
            s1 += A.red[1][t];
        }
        D0b[m] = (float)s0;
        awb[m] = (float)s1;
    }
}

#define CA_ADJ_INST(LPM, NF)                                                                                   \
    template __global__ void k_ensemble_adjoint<LPM, NF>(const DevGeom *, DevNum, const int *, const int64_t *,         \
                                                         const float *, const float *, float *, float *, float *,     \
                                                         float *, float *, EnsTab, float, int, int);
CA_ADJ_INST(64, 33)
CA_ADJ_INST(64, 65)
CA_ADJ_INST(128, 65)
CA_ADJ_INST(64, 129)
CA_ADJ_INST(128, 129)
#undef CA_ADJ_INST
