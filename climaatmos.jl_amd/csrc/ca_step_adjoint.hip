// ca_step_adjoint.hip -- reverse mode (vector-Jacobian products) of the
// sphere step's kernels: the hyperdiffusion Laplacians (k_laplacians), the
// explicit tendency (k_explicit), the implicit Newton stage per column, and
// the elementwise stage algebra.  The transposed weighted DSS is the gather
// DSS kernel run on the transposed tables (ca_dss_table.cpp, transpose).
//
// No reference counterpart: the reference differentiates only in forward
// mode, for its Jacobians (src/prognostic_equations/implicit/autodiff_utils.jl).
// The forward maps reversed here are restated from the same reference files
// as ca_spectral.hip (remaining_tendency.jl:48-58, advection.jl:36-91,
// 205-290, hyperdiffusion.jl:116-147, 273-307,
// vertical_diffusion_boundary_layer.jl:63-103).
//
// Element-local operators: ca_d1 and ca_d1t (ca_d2 and ca_d2t) are each
// other's transposes as maps over the 16 nodes of an element, so the
// transpose of every strong / weak operator is the same DPP lane moves with
// the metric factors applied in reverse order.  Non-smooth points (the
// T_min clamp) take the derivative of the branch the forward took.
#include "ca_kernels.h"

namespace sadj {

// ------------------------------------------------ operator transposes
__device__ __forceinline__ float T_grad(const Nod &q, float o1b, float o2b) { return ca_d1t(q, o1b) + ca_d2t(q, o2b); }
__device__ __forceinline__ void T_div(const Nod &q, const Met &m, float ob, float &v1b, float &v2b) {
    const float t = m.iJ * ob;
    v1b += m.J * ca_d1t(q, t);
    v2b += m.J * ca_d2t(q, t);
}
__device__ __forceinline__ void T_wdiv(const Nod &q, const Met &m, float ob, float &v1b, float &v2b) {
    const float t = -m.iWJ * ob;
    v1b += m.WJ * ca_d1(q, t);
    v2b += m.WJ * ca_d2(q, t);
}
__device__ __forceinline__ float T_wgrad(const Nod &q, const Met &m, float o1b, float o2b) {
    return m.W * (ca_d1(q, -m.iW * o1b) + ca_d2(q, -m.iW * o2b));
}
__device__ __forceinline__ void T_curl3(const Nod &q, const Met &m, float ob, float &u1b, float &u2b) {
    const float t = m.iJ * ob;
    u2b += ca_d1t(q, t);
    u1b -= ca_d2t(q, t);
}
__device__ __forceinline__ void T_wcurl3(const Nod &q, const Met &m, float ob, float &u1b, float &u2b) {
    const float t = m.iWJ * ob;
    u2b -= m.W * ca_d1(q, t);
    u1b += m.W * ca_d2(q, t);
}
__device__ __forceinline__ float T_curl12(const Nod &q, const Met &m, float o1b, float o2b) {
    return ca_d2t(q, m.iJ * o1b) - ca_d1t(q, m.iJ * o2b);
}
__device__ __forceinline__ float T_wcurl12(const Nod &q, const Met &m, float o1b, float o2b) {
    return m.W * (ca_d1(q, m.iWJ * o2b) - ca_d2(q, m.iWJ * o1b));
}
// sum over the 16 nodes of the element (every lane of the DPP row gets it):
// the transpose of ca_elem0 credits it to node (0, 0)
__device__ __forceinline__ float rowsum(float v) {
    v += ca_dpp<CA_ROR(8)>(v);
    v += ca_dpp<CA_ROR(4)>(v);
    v += ca_dpp<CA_ROR(2)>(v);
    v += ca_dpp<CA_ROR(1)>(v);
    return v;
}
// d Pi / d p, d Phi_r / d p (Phi_r(p) = -cp_d [T_min kex ln p + (T_sfc - T_min) Pi^7 / 7]),
// d T_r / d Pi
__device__ __forceinline__ float dexner(const DevNum &n, float p, float Pi) { return n.kex * Pi * ca_rcp(p); }
__device__ __forceinline__ float dphiref(const DevNum &n, float p, float Pi) {
    return -n.cp_d * n.kex * ca_Tref(n, Pi) * ca_rcp(p);
}
__device__ __forceinline__ float dtref(const DevNum &n, float Pi) {
    const float P2 = Pi * Pi, P6 = P2 * P2 * P2;
    return 7.0f * (n.Tsr - n.Tmr) * P6;
}

// ------------------------------------------------ Laplacians (k_laplacians)
// Given Lb, the cotangent of k_laplacians' output at (node, level k), adds
// the centre-state cotangents of the same (node, level) to Yb and the u_3
// cotangents of faces k and k + 1 to fa[k], fb[k] (nothing at the boundary
// faces, which the forward reads as 0).
__global__ void __launch_bounds__(256) k_laplacians_vjp(DevGeom g, DevNum n, const float *Y, const float *Lb,
                                                        float *Yb, float *fa, float *fb) {
    const int col = blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y;
    if (col >= g.ncol) return;
    const Nod q = ca_node(threadIdx.x & 63);
    const int nl = g.nlev;
    const int64_t nC = g.nC;
    const int64_t idx = (int64_t)k * g.ncol + col;
    const float *u3 = Y + 4 * nC;
    const float rho = Y[idx], u1 = Y[nC + idx], u2 = Y[2 * nC + idx], rhoe = Y[3 * nC + idx];
    const float u3a = (k <= 0) ? 0.0f : u3[idx], u3b = (k + 1 >= nl) ? 0.0f : u3[idx + g.ncol];
    const float s = g.sc[k], s2 = s * s, is2 = g.is2c[k], Jv = g.Jvc[k];
    const Met m = ca_met(c_W[q.j * 4 + q.i], c_iW[q.j * 4 + q.i], g.Jh[col] * s2 * Jv, g.iJh[col] * g.isJc[k]);
    float ct1, ct2;
    ca_contra(g, col, is2, u1, u2, ct1, ct2);
    const float ga = g.g33f[k], gb = g.g33f[k + 1];
    const float K = 0.5f * ((u1 * ct1 + u2 * ct2) + 0.5f * (u3a * ga * u3a + u3b * gb * u3b));
    const double e_tot_d = (double)rhoe * ca_drcp((double)rho);
    const double Tu = n.dT_0 + (e_tot_d - (double)K - g.Phic_d[k] + n.dR_d * n.dT_0) * n.dicv_d;
    const bool free_T = Tu > n.dTmin;
    const double Td = fmax(Tu, n.dTmin);
    const float T = (float)Td;
    const float p = (float)((double)rho * n.dR_d * Td);
    const float Pi = ca_exner(n, p);
    const float p0 = ca_elem0(p), Pi0 = ca_elem0(Pi);
    const bool node0 = (q.i | q.j) == 0;
    const float L0b = Lb[idx], L1b = Lb[nC + idx], L2b = Lb[2 * nC + idx], L3b = Lb[3 * nC + idx];
    // vector Laplacian: L_1,2 = wgrad(div c_u) - C(wcurl C(curl u)), L_3 = -Jv^2 wcurl3(...)
    float w1b, w2b;
    ca_cov(g, col, s2, -L0b, -L1b, w1b, w2b);
    const float w3b = Jv * Jv * (-L2b);
    const float cc3b = T_wcurl12(q, m, w1b, w2b);
    float cc1b = 0.0f, cc2b = 0.0f;
    T_wcurl3(q, m, w3b, cc1b, cc2b);
    const float cr3b = Jv * Jv * cc3b;
    float cr1b, cr2b;
    ca_cov(g, col, s2, cc1b, cc2b, cr1b, cr2b);
    const float cu3b = T_curl12(q, m, cr1b, cr2b);
    float u1b = 0.0f, u2b = 0.0f;
    T_curl3(q, m, cr3b, u1b, u2b);
    const float dvb = T_wgrad(q, m, L0b, L1b);
    float ct1b = 0.0f, ct2b = 0.0f;
    T_div(q, m, dvb, ct1b, ct2b);
    // scalar Laplacian of psi = cp_d (T - T_r(Pi)) + Phi_r(p0) - Phi_r(p)
    float v1b = 0.0f, v2b = 0.0f;
    T_wdiv(q, m, L3b, v1b, v2b);
    float q1b, q2b;
    ca_contra(g, col, is2, v1b, v2b, q1b, q2b);
    const float psib = T_grad(q, q1b, q2b);
    const float Tb0 = n.cp_d * psib;
    const float Pib = -n.cp_d * dtref(n, Pi) * psib;
    float pb = -dphiref(n, p, Pi) * psib;
    const float sp = rowsum(psib);
    if (node0) pb += dphiref(n, p0, Pi0) * sp;
    pb += Pib * dexner(n, p, Pi);
    // p = rho R_d T, T = max(T_u, T_min), T_u = T_0 + (e_tot - K - Phi + R_d T_0) / cv_d
    float rhob = pb * n.R_d * T;
    const float Tdb = Tb0 + pb * rho * n.R_d;
    const float Tub = free_T ? Tdb : 0.0f;
    const float icv = (float)n.dicv_d;
    const float etb = Tub * icv, Kb = -Tub * icv;
    const float irho = ca_rcp(rho);
    const float rhoeb = etb * irho;
    rhob -= etb * rhoe * irho * irho;
    u1b += Kb * ct1;
    u2b += Kb * ct2;
    float a1, a2;
    ca_contra(g, col, is2, ct1b, ct2b, a1, a2);
    u1b += a1;
    u2b += a2;
    const float u3ab = Kb * 0.5f * ga * u3a + 0.5f * cu3b;
    const float u3bb = Kb * 0.5f * gb * u3b + 0.5f * cu3b;
    Yb[idx] += rhob;
    Yb[nC + idx] += u1b;
    Yb[2 * nC + idx] += u2b;
    Yb[3 * nC + idx] += rhoeb;
    if (k >= 1) fa[idx] += u3ab;
    if (k + 1 <= nl - 1) fb[idx] += u3bb;
}

// ------------------------------------------------ explicit tendency (k_explicit)
// Forward of one centre of k_explicit (sphere) with every intermediate its
// reverse needs.  The arithmetic is k_explicit's.
struct Loc {
    float rho, u1, u2, rhoe, u3a, u3b;
    float s, s2, is2, Jv;
    Met m;
    float ct1, ct2, Kh, Kv, T, h, p, Pi;
    bool free_T;
    float m1, m2, wd, h0, hh, gh1, gh2, p0, Pi0, a1, a2, c1, c2, th, thPi;
    float s1, s2g;  // grad of the scalar Laplacian (hyperdiffusion)
    float a3;
    float dr, d1, d2, de;                              // local tendency parts
    float rhoJ, irhoJ, l4, sd, uc0, uc1, uc2, irho;   // values the faces read
};

template <bool HYPER, bool VDIFF>
__device__ __forceinline__ void loc_forward(Loc &c, const DevGeom &g, const DevNum &n, const Nod &q, float W, float iW,
                                            int col, int k, const float *Y, const float *L, const ColMet &cm) {
    const int64_t nC = g.nC, nc = g.ncol;
    const int nl = g.nlev;
    const uint32_t cb = (uint32_t)col * 4u;  // the lane's byte offset in every level row
    const float *Yk = Y + (int64_t)k * nc;
    c.rho = ca_ldo(Yk, cb);
    c.u1 = ca_ldo(Yk + nC, cb);
    c.u2 = ca_ldo(Yk + 2 * nC, cb);
    c.rhoe = ca_ldo(Yk + 3 * nC, cb);
    c.u3a = k <= 0 ? 0.0f : ca_ldo(Yk + 4 * nC, cb);
    c.u3b = k + 1 >= nl ? 0.0f : ca_ldo(Yk + 4 * nC + nc, cb);
    c.s = g.sc[k];
    c.s2 = c.s * c.s;
    c.is2 = g.is2c[k];
    c.Jv = g.Jvc[k];
    c.m = ca_met(W, iW, g.Jh[col] * c.s2 * c.Jv, g.iJh[col] * g.isJc[k]);
    const Met &m = c.m;
    ca_contra(cm, c.is2, c.u1, c.u2, c.ct1, c.ct2);
    c.Kh = 0.5f * (c.u1 * c.ct1 + c.u2 * c.ct2);
    c.Kv = 0.25f * (c.u3a * g.g33f[k] * c.u3a + c.u3b * g.g33f[k + 1] * c.u3b);
    const float K = c.Kh + c.Kv;
    const double e_tot_d = (double)c.rhoe * ca_drcp((double)c.rho);
    const double Tu = n.dT_0 + (e_tot_d - (double)K - g.Phic_d[k] + n.dR_d * n.dT_0) * n.dicv_d;
    c.free_T = Tu > n.dTmin;
    const double Td = fmax(Tu, n.dTmin);
    c.T = (float)Td;
    c.h = (float)(e_tot_d + n.dR_d * Td);
    c.p = (float)((double)c.rho * n.dR_d * Td);
    c.Pi = ca_exner(n, c.p);
    float dr = 0.0f, d1 = 0.0f, d2 = 0.0f, de = 0.0f;
    // horizontal dynamics
    c.m1 = c.rho * c.ct1;
    c.m2 = c.rho * c.ct2;
    c.wd = op_wdiv(q, m, c.m1, c.m2);
    dr -= 0.5f * c.wd + 0.5f * c.wd;
    c.h0 = ca_elem0(c.h);
    c.hh = c.h - c.h0;
    const float wdh = op_wdiv(q, m, c.m1 * c.hh, c.m2 * c.hh);
    op_grad(q, c.hh, c.gh1, c.gh2);
    de -= 0.5f * wdh + 0.5f * (c.hh * c.wd + (c.m1 * c.gh1 + c.m2 * c.gh2)) + c.h0 * c.wd;
    c.p0 = ca_elem0(c.p);
    c.Pi0 = ca_elem0(c.Pi);
    const float ip0 = ca_rcp(c.p0);
    float dPhr, dPi;
    ca_dref(n, c.p0, ip0, c.Pi0, c.p, dPhr, dPi);
    const float A = K - dPhr;
    c.thPi = c.T - ca_Tref(n, c.Pi);
    c.th = c.thPi * ca_rcp(c.Pi);
    float g1, g2, b1, b2;
    op_grad(q, A, g1, g2);
    op_grad(q, dPi, c.a1, c.a2);
    op_grad(q, c.thPi, b1, b2);
    op_grad(q, c.th, c.c1, c.c2);
    d1 -= g1 + n.cp_d * (c.th * c.a1 + b1 - c.Pi * c.c1) * 0.5f;
    d2 -= g2 + n.cp_d * (c.th * c.a2 + b2 - c.Pi * c.c2) * 0.5f;
    c.l4 = 0.0f;
    c.s1 = c.s2g = 0.0f;
    if (HYPER) {
        const float *Lk = L + (int64_t)k * nc;
        const float X1 = ca_ldo(Lk, cb), X2 = ca_ldo(Lk + nC, cb), X3 = ca_ldo(Lk + 2 * nC, cb),
                    Ls = ca_ldo(Lk + 3 * nC, cb);
        float v1, v2;
        ca_contra(cm, c.is2, X1, X2, v1, v2);
        const float dv = op_div(q, m, v1, v2);
        float gd1, gd2;
        op_wgrad(q, m, dv, gd1, gd2);
        const float cr3 = op_curl3(q, m, X1, X2);
        float cr1, cr2;
        op_curl12(q, m, X3, cr1, cr2);
        float cc1, cc2;
        ca_cov(cm, c.s2, cr1, cr2, cc1, cc2);
        const float cc3 = c.Jv * c.Jv * cr3;
        const float w3 = op_wcurl3(q, m, cc1, cc2);
        float w1, w2;
        op_wcurl12(q, m, cc3, w1, w2);
        float wc1, wc2;
        ca_cov(cm, c.s2, w1, w2, wc1, wc2);
        const float wc3 = c.Jv * c.Jv * w3;
        d1 -= n.nu_v * (n.dd * gd1 - wc1);
        d2 -= n.nu_v * (n.dd * gd2 - wc2);
        c.l4 = n.dd * 0.0f - wc3;
        op_grad(q, Ls, c.s1, c.s2g);
        float r1, r2;
        ca_contra(cm, c.is2, c.rho * c.s1, c.rho * c.s2g, r1, r2);
        de -= n.nu_s * op_wdiv(q, m, r1, r2);
    }
    const float om3 = op_wcurl3(q, m, c.u1, c.u2);
    c.a3 = g.f3n[col] * g.iJvc[k] + om3;
    d1 -= m.J * (0.0f - c.a3 * c.ct2);
    d2 -= m.J * (c.a3 * c.ct1 - 0.0f);
    if (n.rayleigh && n.alpha_uh != 0.0f) {
        const float beta = n.alpha_uh * g.zeta_c[k];
        d1 += -beta * c.u1;
        d2 += -beta * c.u2;
    }
    c.dr = dr;
    c.d1 = d1;
    c.d2 = d2;
    c.de = de;
    c.rhoJ = c.rho * m.J;
    c.irhoJ = ca_rcp(c.rhoJ);
    c.irho = ca_rcp(c.rho);
    c.sd = c.uc0 = c.uc1 = c.uc2 = 0.0f;
    if (VDIFF) {
        c.sd = n.cp_d * (c.T - n.T_0) + g.Phic[k];
        const float *B = g.basis;
        const float is = ca_rcp(c.s);
        c.uc0 = (c.u1 * B[0 * nc + col] + c.u2 * B[3 * nc + col]) * is;
        c.uc1 = (c.u1 * B[1 * nc + col] + c.u2 * B[4 * nc + col]) * is;
        c.uc2 = (c.u1 * B[2 * nc + col] + c.u2 * B[5 * nc + col]) * is;
    }
}

// cotangents of the values a centre hands to its faces and completion
struct LocBar {
    float rhoJ, irhoJ, Kh, Kv, ct1, ct2, u1, u2, rho, irho, l4, sd, uc0, uc1, uc2;
};
__device__ __forceinline__ void lb_zero(LocBar &b) {
    b.rhoJ = b.irhoJ = b.Kh = b.Kv = b.ct1 = b.ct2 = b.u1 = b.u2 = b.rho = b.irho = b.l4 = b.sd = 0.0f;
    b.uc0 = b.uc1 = b.uc2 = 0.0f;
}

// Reverse of loc_forward: given the cotangents of the local tendency parts
// (db: dr, d1, d2, de) and of the face values (vb), returns the cotangents
// of this node's rho, u1, u2, rho e_tot (ub[0..3]), of the u_3 of faces k,
// k + 1 (u3ab, u3bb) and of the four Laplacian planes (Lb).  Every lane of
// the element takes part (DPP); d0b: the D0 cotangent is not touched here.
template <bool HYPER, bool VDIFF>
__device__ __forceinline__ void loc_reverse(const Loc &c, const LocBar &vb, const float *db, const DevGeom &g,
                                            const DevNum &n, const Nod &q, int col, int k, float *ub, float &u3ab,
                                            float &u3bb, float *Lb, const ColMet &cm) {
    const Met &m = c.m;
    const int64_t nc = g.ncol;
    const bool node0 = (q.i | q.j) == 0;
    float rhob = 0.0f, u1b = 0.0f, u2b = 0.0f, Tb = 0.0f, Khb = vb.Kh, Kvb = vb.Kv, ct1b = vb.ct1, ct2b = vb.ct2;
    // face values
    const float rhoJb = vb.rhoJ - c.irhoJ * c.irhoJ * vb.irhoJ;
    rhob += m.J * rhoJb + vb.rho - c.irho * c.irho * vb.irho;
    u1b += vb.u1;
    u2b += vb.u2;
    if (VDIFF) {
        Tb += n.cp_d * vb.sd;
        const float *B = g.basis;
        const float is = ca_rcp(c.s);
        u1b += (vb.uc0 * B[0 * nc + col] + vb.uc1 * B[1 * nc + col] + vb.uc2 * B[2 * nc + col]) * is;
        u2b += (vb.uc0 * B[3 * nc + col] + vb.uc1 * B[4 * nc + col] + vb.uc2 * B[5 * nc + col]) * is;
    }
    const float drb = db[0], d1b = db[1], d2b = db[2], deb = db[3];
    // Rayleigh u_h sponge
    if (n.rayleigh && n.alpha_uh != 0.0f) {
        const float beta = n.alpha_uh * g.zeta_c[k];
        u1b -= beta * d1b;
        u2b -= beta * d2b;
    }
    // vertical advection, centre part: d1 += J a3 ct2, d2 -= J a3 ct1, a3 = f3 / Jv + wcurl3(u)
    {
        const float a3b = m.J * (c.ct2 * d1b - c.ct1 * d2b);
        ct2b += m.J * c.a3 * d1b;
        ct1b -= m.J * c.a3 * d2b;
        T_wcurl3(q, m, a3b, u1b, u2b);
    }
    // hyperdiffusion
    if (HYPER) {
        float r1b = 0.0f, r2b = 0.0f;
        T_wdiv(q, m, -n.nu_s * deb, r1b, r2b);
        float rs1b, rs2b;
        ca_contra(cm, c.is2, r1b, r2b, rs1b, rs2b);
        rhob += c.s1 * rs1b + c.s2g * rs2b;
        Lb[3] = T_grad(q, c.rho * rs1b, c.rho * rs2b);
        const float gd1b = -n.nu_v * n.dd * d1b, gd2b = -n.nu_v * n.dd * d2b;
        float w1b, w2b;
        ca_cov(cm, c.s2, n.nu_v * d1b, n.nu_v * d2b, w1b, w2b);
        const float w3b = c.Jv * c.Jv * (-vb.l4);
        const float cc3b = T_wcurl12(q, m, w1b, w2b);
        float cc1b = 0.0f, cc2b = 0.0f;
        T_wcurl3(q, m, w3b, cc1b, cc2b);
        const float cr3b = c.Jv * c.Jv * cc3b;
        float cr1b, cr2b;
        ca_cov(cm, c.s2, cc1b, cc2b, cr1b, cr2b);
        Lb[2] = T_curl12(q, m, cr1b, cr2b);
        float X1b = 0.0f, X2b = 0.0f;
        T_curl3(q, m, cr3b, X1b, X2b);
        const float dvb = T_wgrad(q, m, gd1b, gd2b);
        float v1b = 0.0f, v2b = 0.0f;
        T_div(q, m, dvb, v1b, v2b);
        float xa, xb;
        ca_contra(cm, c.is2, v1b, v2b, xa, xb);
        Lb[0] = X1b + xa;
        Lb[1] = X2b + xb;
    } else {
        Lb[0] = Lb[1] = Lb[2] = Lb[3] = 0.0f;
    }
    // pressure gradient: d = -(grad A + cp (th grad dPi + grad thPi - Pi grad th) / 2)
    float Pib = 0.5f * n.cp_d * (c.c1 * d1b + c.c2 * d2b);
    float thb = -0.5f * n.cp_d * (c.a1 * d1b + c.a2 * d2b);
    const float Ab = T_grad(q, -d1b, -d2b);
    const float dPib = T_grad(q, -0.5f * n.cp_d * c.th * d1b, -0.5f * n.cp_d * c.th * d2b);
    float thPib = T_grad(q, -0.5f * n.cp_d * d1b, -0.5f * n.cp_d * d2b);
    thb += T_grad(q, 0.5f * n.cp_d * c.Pi * d1b, 0.5f * n.cp_d * c.Pi * d2b);
    const float iPi = ca_rcp(c.Pi);
    thPib += thb * iPi;
    Pib -= thb * c.th * iPi;
    Tb += thPib;
    Pib -= thPib * dtref(n, c.Pi);
    // dPi = Pi(p) - Pi(p0); A = K - (Phi_r(p) - Phi_r(p0))
    float pb = dPib * dexner(n, c.p, c.Pi) - Ab * dphiref(n, c.p, c.Pi);
    float p0b = -dPib * dexner(n, c.p0, c.Pi0) + Ab * dphiref(n, c.p0, c.Pi0);
    float Kb = Ab;
    // energy flux: de = -(wdiv(m hh)/2 + (hh wd + m . grad hh)/2 + h0 wd), dr = -wd
    float wdb = -(0.5f * c.hh + c.h0) * deb - drb;
    float hhb = -0.5f * c.wd * deb;
    float m1b = -0.5f * c.gh1 * deb, m2b = -0.5f * c.gh2 * deb;
    float h0b = -c.wd * deb;
    hhb += T_grad(q, -0.5f * c.m1 * deb, -0.5f * c.m2 * deb);
    {
        float mh1b = 0.0f, mh2b = 0.0f;
        T_wdiv(q, m, -0.5f * deb, mh1b, mh2b);
        m1b += c.hh * mh1b;
        m2b += c.hh * mh2b;
        hhb += c.m1 * mh1b + c.m2 * mh2b;
    }
    float hb = hhb;
    h0b -= hhb;
    T_wdiv(q, m, wdb, m1b, m2b);
    rhob += c.ct1 * m1b + c.ct2 * m2b;
    ct1b += c.rho * m1b;
    ct2b += c.rho * m2b;
    // element-first-node values (ca_elem0) credit node (0, 0)
    const float P0 = rowsum(p0b), H0 = rowsum(h0b);
    if (node0) {
        pb += P0;
        hb += H0;
    }
    pb += Pib * dexner(n, c.p, c.Pi);
    // thermodynamics: h = e_tot + R T, p = rho R T, T = max(T_u, T_min)
    const float Tdb = Tb + n.R_d * hb + c.rho * n.R_d * pb;
    rhob += n.R_d * c.T * pb;
    const float Tub = c.free_T ? Tdb : 0.0f;
    const float icv = (float)n.dicv_d;
    const float etb = hb + icv * Tub;
    Kb -= icv * Tub;
    const float rhoeb = etb * c.irho;
    rhob -= etb * c.rhoe * c.irho * c.irho;
    Khb += Kb;
    Kvb += Kb;
    u1b += Khb * c.ct1;
    u2b += Khb * c.ct2;
    u3ab = Kvb * 0.5f * g.g33f[k] * c.u3a;
    u3bb = Kvb * 0.5f * g.g33f[k + 1] * c.u3b;
    float a1, a2;
    ca_contra(cm, c.is2, ct1b, ct2b, a1, a2);
    ub[0] = rhob;
    ub[1] = u1b + a1;
    ub[2] = u2b + a2;
    ub[3] = rhoeb;
}

// Face f forward values the centre completions read.
struct FaceF {
    float G1, G2, Fs, Fu0, Fu1, Fu2;
};

// Reverse of k_explicit (sphere).  Ytb: cotangent of the tendency; Y, L:
// the state and the DSSed Laplacians it was evaluated at.  Writes the state
// cotangent Yb (centre planes of the owned centres, u_3 of the owned faces),
// the cotangent Lb of the Laplacians (owned centres), and adds the D0
// cotangent of the owned faces to d0acc[chunk * ncol + col].  The u_3
// cotangent is split: Yb's face plane gets each face's own terms, and the
// local reverse of centre k writes its u_3 terms for faces k and k + 1 to
// fa[k], fb[k] (zero on the boundary faces), which the caller adds.
// Work item = (node, chunk of levels) as k_explicit; chunk c owns centres
// and faces [k0, k1) (and face nlev when k1 = nlev).  The sweep goes up from
// centre k0 - 1: face f's reverse gives the centres f - 1 and f their face
// cotangents, after which centre f - 1 is complete and, if owned, its local
// reverse runs.
template <bool HYPER, bool VDIFF>
__global__ void __launch_bounds__(256) k_explicit_vjp(DevGeom g, DevNum n, const float *Y, const float *L,
                                                      const float *Ytb, float *Yb, float *Lb, float *fa, float *fb,
                                                      double *d0acc, int chunk) {
    const int col = blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= g.ncol) return;
    const Nod q = ca_node(threadIdx.x & 63);
    const int nl = g.nlev;
    const int64_t nC = g.nC, nc = g.ncol;
    const float *u3 = Y + 4 * nC;
    const float W = c_W[q.j * 4 + q.i], iW = c_iW[q.j * 4 + q.i];
    const float Jh = g.Jh[col], iJh = g.iJh[col];
    const float f12a = g.f12a[col], f12b = g.f12b[col];
    const ColMet cm = ca_colmet(g, col);
    const uint32_t cbo = (uint32_t)col * 4u;  // the lane's byte offset in every level row
    float Kf = 0.0f;
    bool kfree = false;
    if (VDIFF) {
        const float D0 = g.D0_col[col];
        const float Kh = fmaxf(D0, 1.1920929e-07f);
        Kf = 1.0f / (0.5f * (1.0f / Kh + 1.0f / Kh));
        kfree = D0 > 1.1920929e-07f;
    }
    const int k0 = chunk > 0 ? (int)blockIdx.y * chunk : 0;
    const int k1 = chunk > 0 ? min(k0 + chunk, nl) : nl;
    const int kbeg = max(k0 - 1, 0);
    Loc prev, cur;
    LocBar pb, cb;
    lb_zero(pb);
    lb_zero(cb);
    float pY[4] = {0, 0, 0, 0}, cY[4] = {0, 0, 0, 0};  // tendency cotangents of prev / cur centres
    FaceF fprev{0, 0, 0, 0, 0, 0};                      // face below prev
    double d0 = 0.0;
    for (int k = kbeg; k <= k1; ++k) {
        // ---- centre k: forward, tendency cotangent
        if (k < nl) {
            loc_forward<HYPER, VDIFF>(cur, g, n, q, W, iW, col, k, Y, L, cm);
            const float *Tk = Ytb + (int64_t)k * nc;
            cY[0] = ca_ldo(Tk, cbo);
            cY[1] = ca_ldo(Tk + nC, cbo);
            cY[2] = ca_ldo(Tk + 2 * nC, cbo);
            cY[3] = ca_ldo(Tk + 3 * nC, cbo);
        }
        lb_zero(cb);
        // ---- face k: forward and reverse (interior faces, prev = k - 1 valid)
        FaceF fk{0, 0, 0, 0, 0, 0};
        float u3face = 0.0f;
        const bool interior = k >= 1 && k <= nl - 1 && k > kbeg;
        {
            const float sf = g.sf[k];
            const Met mf = ca_met(W, iW, Jh * sf * sf * g.Jvf[k], iJh * g.isJf[k]);
            const float Jf = mf.J, iJf = mf.iJ;
            const float u3f = (k >= 1 && k <= nl - 1) ? ca_ldo(u3 + (int64_t)k * nc, cbo) : 0.0f;
            float x = 0.0f, y = 0.0f;
            if (interior) op_wcurl12(q, mf, u3f, x, y);
            if (interior) {
                const float fu3 = u3f * g.g33f[k];
                const float wa = -(cur.u2 - prev.u2) * iJf + x;
                const float wb = (cur.u1 - prev.u1) * iJf + y;
                const float isf = g.isf[k];
                const float A1 = f12a * isf + wa, A2 = f12b * isf + wb;
                const float rJf = 0.5f * (prev.rhoJ + cur.rhoJ);
                const float b3 = rJf * fu3;
                fk.G1 = Jf * (A2 * b3);
                fk.G2 = Jf * (0.0f - A1 * b3);
                const float i1 = 0.5f * (prev.ct1 + cur.ct1), i2 = 0.5f * (prev.ct2 + cur.ct2);
                float rK = 0.0f, iJv = 0.0f;
                if (VDIFF) {
                    rK = (0.5f * (prev.rho + cur.rho)) * Kf;
                    iJv = g.iJvf[k];
                    fk.Fs = Jf * (iJv * iJv * (-(rK * (cur.sd - prev.sd))));
                    const float fac = Jf * iJv * (-2.0f * rK) * 0.5f;
                    fk.Fu0 = fac * (cur.uc0 - prev.uc0) * iJv;
                    fk.Fu1 = fac * (cur.uc1 - prev.uc1) * iJv;
                    fk.Fu2 = fac * (cur.uc2 - prev.uc2) * iJv;
                }
                // cotangents of this face's outputs, from the completions of centres k - 1 and k
                const float du3b = ca_ldo(Ytb + 4 * nC + (int64_t)k * nc, cbo);
                const float G1b = -0.5f * (prev.irhoJ * pY[1] + cur.irhoJ * cY[1]);
                const float G2b = -0.5f * (prev.irhoJ * pY[2] + cur.irhoJ * cY[2]);
                // du3 = -(Jf (A1 i2 - A2 i1) + dKh + dKv) - nu_v <l4>_rhoJ
                float A1b = -Jf * i2 * du3b, A2b = Jf * i1 * du3b;
                const float i1b = Jf * A2 * du3b, i2b = -Jf * A1 * du3b;
                cb.Kh -= du3b;
                pb.Kh += du3b;
                cb.Kv -= du3b;
                pb.Kv += du3b;
                pb.ct1 += 0.5f * i1b;
                cb.ct1 += 0.5f * i1b;
                pb.ct2 += 0.5f * i2b;
                cb.ct2 += 0.5f * i2b;
                if (HYPER) {
                    const float wA = prev.rhoJ, wB = cur.rhoJ;
                    const float den = 0.5f * (wA + wB), iden = ca_rcp(den);
                    const float num = 0.5f * (wA * prev.l4 + wB * cur.l4);
                    const float qb = -n.nu_v * du3b;
                    const float numb = qb * iden, denb = -qb * num * iden * iden;
                    pb.l4 += 0.5f * wA * numb;
                    cb.l4 += 0.5f * wB * numb;
                    pb.rhoJ += 0.5f * prev.l4 * numb + 0.5f * denb;
                    cb.rhoJ += 0.5f * cur.l4 * numb + 0.5f * denb;
                }
                // G1 = Jf A2 b3, G2 = -Jf A1 b3, b3 = rJf u3 g33
                A2b += Jf * b3 * G1b;
                A1b -= Jf * b3 * G2b;
                const float b3b = Jf * (A2 * G1b - A1 * G2b);
                const float rJfb = fu3 * b3b;
                u3face += g.g33f[k] * rJf * b3b;
                pb.rhoJ += 0.5f * rJfb;
                cb.rhoJ += 0.5f * rJfb;
                // wa = -(u2_k - u2_k-1) / Jf + x, wb = (u1_k - u1_k-1) / Jf + y
                cb.u2 -= iJf * A1b;
                pb.u2 += iJf * A1b;
                cb.u1 += iJf * A2b;
                pb.u1 -= iJf * A2b;
                u3face += T_wcurl12(q, mf, A1b, A2b);
                if (VDIFF) {
                    const float iJca = prev.m.iJ, iJcb = cur.m.iJ;
                    const float Fsb = -iJca * pY[3] + iJcb * cY[3];
                    // e_i of centre c = (Fu_i(top) - Fu_i(bottom)) iJc; c1,2 = s (e . B6.., B9..); d1,2 -= c / rho
                    const float *B = g.basis;
                    const float pc1b = -prev.irho * pY[1], pc2b = -prev.irho * pY[2];
                    const float cc1b = -cur.irho * cY[1], cc2b = -cur.irho * cY[2];
                    float Fub[3];
#pragma unroll
                    for (int i = 0; i < 3; ++i) {
                        const float Bi6 = B[(6 + i) * nc + col], Bi9 = B[(9 + i) * nc + col];
                        const float peb = prev.s * (Bi6 * pc1b + Bi9 * pc2b);
                        const float ceb = cur.s * (Bi6 * cc1b + Bi9 * cc2b);
                        Fub[i] = iJca * peb - iJcb * ceb;
                    }
                    const float c0 = Jf * iJv * iJv;
                    float rKb = -c0 * (cur.sd - prev.sd) * Fsb;
                    cb.sd -= c0 * rK * Fsb;
                    pb.sd += c0 * rK * Fsb;
                    rKb -= c0 * ((cur.uc0 - prev.uc0) * Fub[0] + (cur.uc1 - prev.uc1) * Fub[1] +
                                 (cur.uc2 - prev.uc2) * Fub[2]);
                    cb.uc0 -= c0 * rK * Fub[0];
                    pb.uc0 += c0 * rK * Fub[0];
                    cb.uc1 -= c0 * rK * Fub[1];
                    pb.uc1 += c0 * rK * Fub[1];
                    cb.uc2 -= c0 * rK * Fub[2];
                    pb.uc2 += c0 * rK * Fub[2];
                    pb.rho += 0.5f * Kf * rKb;
                    cb.rho += 0.5f * Kf * rKb;
                    if (kfree && k >= k0 && k < k1) d0 += (double)(0.5f * (prev.rho + cur.rho) * rKb);
                }
            }
        }
        if ((k >= k0 && k < k1) || (k == nl && k1 == nl)) ca_sto(Yb + 4 * nC + (int64_t)k * nc, cbo, u3face);
        // ---- owned centre k - 1 complete: completion terms that need both of its faces, then its local reverse
        if (k - 1 >= k0 && k > kbeg) {
            const int kk = k - 1;
            pb.irhoJ += -0.5f * (fprev.G1 + fk.G1) * pY[1] - 0.5f * (fprev.G2 + fk.G2) * pY[2];
            if (VDIFF) {
                const float iJc = prev.m.iJ;
                const float *B = g.basis;
                const float e0 = (fk.Fu0 - fprev.Fu0) * iJc, e1 = (fk.Fu1 - fprev.Fu1) * iJc,
                            e2 = (fk.Fu2 - fprev.Fu2) * iJc;
                const float c1 = prev.s * (e0 * B[6 * nc + col] + e1 * B[7 * nc + col] + e2 * B[8 * nc + col]);
                const float c2 = prev.s * (e0 * B[9 * nc + col] + e1 * B[10 * nc + col] + e2 * B[11 * nc + col]);
                pb.irho += -(c1 * pY[1] + c2 * pY[2]);
            }
            float ubc[4], Lbc[4], u3ab, u3bb;
            loc_reverse<HYPER, VDIFF>(prev, pb, pY, g, n, q, col, kk, ubc, u3ab, u3bb, Lbc, cm);
            const int64_t row = (int64_t)kk * nc;
            ca_sto(Yb + row, cbo, ubc[0]);
            ca_sto(Yb + nC + row, cbo, ubc[1]);
            ca_sto(Yb + 2 * nC + row, cbo, ubc[2]);
            ca_sto(Yb + 3 * nC + row, cbo, ubc[3]);
            ca_sto(Lb + row, cbo, Lbc[0]);
            ca_sto(Lb + nC + row, cbo, Lbc[1]);
            ca_sto(Lb + 2 * nC + row, cbo, Lbc[2]);
            ca_sto(Lb + 3 * nC + row, cbo, Lbc[3]);
            ca_sto(fa + row, cbo, kk >= 1 ? u3ab : 0.0f);
            ca_sto(fb + row, cbo, kk + 1 <= nl - 1 ? u3bb : 0.0f);
        }
        fprev = fk;
        prev = cur;
        pb = cb;
        pY[0] = cY[0];
        pY[1] = cY[1];
        pY[2] = cY[2];
        pY[3] = cY[3];
    }
    if (VDIFF) d0acc[(int64_t)blockIdx.y * nc + col] += d0;
}

template __global__ void k_explicit_vjp<true, true>(DevGeom, DevNum, const float *, const float *, const float *,
                                                    float *, float *, float *, float *, double *, int);
template __global__ void k_explicit_vjp<true, false>(DevGeom, DevNum, const float *, const float *, const float *,
                                                     float *, float *, float *, float *, double *, int);
template __global__ void k_explicit_vjp<false, true>(DevGeom, DevNum, const float *, const float *, const float *,
                                                     float *, float *, float *, float *, double *, int);
template __global__ void k_explicit_vjp<false, false>(DevGeom, DevNum, const float *, const float *, const float *,
                                                      float *, float *, float *, float *, double *, int);

// ------------------------------------------------ implicit stage, per column
// Reverse of one Newton stage (k_newton_pcr) on CPB columns per workgroup,
// LPM lanes each: given U~ (Ut), the stage's recorded result U (Un) and the
// cotangent Ub of U, writes the cotangent of U~ to Utb.  Each column is
// viewed as a one-column grid (cg[col]: the sphere's per-column planes
// offset to the column) and reversed by the column ensemble's
// adj::newton_vjp, which rebuilds the stage's internals from U~ and U.  With LPM = 32 two columns share a wave; a column index past the
// grid's end (odd column counts) repeats the last column without writing, so
// every lane reaches every barrier.
template <int LPM, int NF, int CPB>
__global__ void __launch_bounds__(LPM * CPB) __attribute__((amdgpu_waves_per_eu(4)))
k_newton_vjp(const DevGeom *__restrict__ cg, DevNum n, int64_t nC, int64_t nc, const float *__restrict__ Ut,
             const float *__restrict__ Un, const float *__restrict__ Ub, float dtg, float *__restrict__ Utb,
             const float *__restrict__ fib, float cfi, const float *__restrict__ fa, const float *__restrict__ fb) {
    using namespace adj;
    __shared__ NewtonAdjShared<NF> As[CPB];
    const int sub = CPB == 1 ? 0 : (int)threadIdx.x / LPM;
    const int tid = CPB == 1 ? (int)threadIdx.x : (int)threadIdx.x % LPM;
    const int64_t cwant = (int64_t)ca_xcd_tile(blockIdx.x, gridDim.x) * CPB + sub;
    const bool live = cwant < nc;
    const int64_t col = live ? cwant : nc - 1;
    NewtonAdjShared<NF> &A = As[sub];
    const DevGeom &g = cg[col];
    const int nl = g.nlev;
    // the cotangent of U: Ub, + F_I's direct term -cfi fib (rho, rho e_tot,
    // u_3) and the Laplacian reverse's face terms fa, fb when given
    for (int k = tid; k < nl; k += LPM)
        for (int p = 0; p < 4; ++p) {
            const int64_t i = p * nC + (int64_t)k * nc + col;
            const float v = Ut[i];
            A.S.U[p][k] = v;
            A.Ut[0][p][k] = v;
            A.Ub[p][k] = (p == 0 || p == 3) ? fmaf(-cfi, fib[i], Ub[i]) : Ub[i];
            if (p == 0) A.W[22][k] = Un[i];
        }
    for (int f = tid; f <= nl; f += LPM) {
        const bool in = f >= 1 && f <= nl - 1;
        const int64_t q = (int64_t)f * nc + col, i = 4 * nC + q;
        const float v = in ? Ut[i] : 0.0f;
        A.S.U[4][f] = v;
        A.Ut[0][4][f] = v;
        A.W[23][f] = in ? Un[i] : 0.0f;
        float b = in ? fmaf(-cfi, fib[i], Ub[i]) : 0.0f;
        if (in && fa) b += fa[q] + fb[q - nc];
        A.Ub[4][f] = b;
    }
    __syncthreads();
    (void)newton_vjp<LPM>(A, g, n, dtg, A.Ut[0], A.W[22], A.W[23], tid);  // W[22], W[23]: rho, u_3 of U
    if (!live) return;
    // U~ also enters F_I = (U - U~) / dtg directly: Utb += cfi fib on rho, rho e_tot, u_3
    for (int k = tid; k < nl; k += LPM)
        for (int p = 0; p < 4; ++p) {
            const int64_t i = p * nC + (int64_t)k * nc + col;
            Utb[i] = (p == 0 || p == 3) ? fmaf(cfi, fib[i], A.Utb[p][k]) : A.Utb[p][k];
        }
    for (int f = tid; f <= nl; f += LPM) {
        const int64_t i = 4 * nC + (int64_t)f * nc + col;
        Utb[i] = fmaf(cfi, fib[i], A.Utb[4][f]);
    }
}

#define CA_NVJP_INST(LPM, NF, CPB)                                                                                   \
    template __global__ void k_newton_vjp<LPM, NF, CPB>(const DevGeom *, DevNum, int64_t, int64_t, const float *,   \
                                                        const float *, const float *, float, float *,              \
                                                        const float *, float, const float *, const float *);
CA_NVJP_INST(32, 33, 2)
CA_NVJP_INST(64, 65, 1)
CA_NVJP_INST(128, 129, 1)
#undef CA_NVJP_INST

// ------------------------------------------------ stage algebra
// U~ of an implicit stage from its sources (the assembly of k_newton_pcr),
// with u_3 = 0 on the boundary faces.  Thread per (column, level).
__global__ void __launch_bounds__(256) k_stage_state(DevGeom g, StageSrc src, float *out) {
    const int col = blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y;
    if (col >= g.ncol) return;
    const int nl = g.nlev;
    const int64_t q = (int64_t)k * g.ncol + col;
    float v[5];
    assemble<-1, -1>(src, g.nC, q, true, k < nl, k >= 1 && k <= nl - 1, v);
    if (k < nl)
        for (int p = 0; p < 4; ++p) out[p * g.nC + q] = v[p];
    out[4 * g.nC + q] = v[4];
}

// out = (acc ? out : 0) + sum_t a_t x_t over the planes selected by the
// bit mask (bit p: plane p, 0..3 centres, 4 faces); with fa, fb, the
// interior faces of out also get fa[f] + fb[f - 1] (the Laplacian reverse's
// face terms); with out2, out2 = sum_t a2_t x_t over planes2 as well.
struct Lin {
    const float *x[6];
    float a[6], a2[6];
    int nx, acc;
    unsigned planes, planes2;
    float *out2;
    const float *fa, *fb;
};
__global__ void __launch_bounds__(256) k_lincomb(DevGeom g, float *out, Lin L) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t S = 4 * g.nC + g.nF;
    if (i >= S) return;
    const int p = i < 4 * g.nC ? (int)(i / g.nC) : 4;
    float x[6];
    for (int t = 0; t < L.nx; ++t) x[t] = L.x[t][i];
    if ((L.planes >> p) & 1u) {
        float v = L.acc ? out[i] : 0.0f;
        for (int t = 0; t < L.nx; ++t) v = fmaf(L.a[t], x[t], v);
        if (p == 4 && L.fa) {
            const int64_t q = i - 4 * g.nC;
            const int64_t f = q / g.ncol;
            if (f >= 1 && f <= g.nlev - 1) v += L.fa[q] + L.fb[q - g.ncol];
        }
        out[i] = v;
    }
    if (L.out2 && ((L.planes2 >> p) & 1u)) {
        float v = 0.0f;
        for (int t = 0; t < L.nx; ++t) v = fmaf(L.a2[t], x[t], v);
        L.out2[i] = v;
    }
}

// As k_lincomb, 4 consecutive entries per thread with 16-byte accesses (every
// plane a multiple of 4 entries and every pointer 16-byte aligned: the host
// checks); the same arithmetic per entry.
__device__ __forceinline__ float4 lin_ld4(const float *p, int64_t i) { return *reinterpret_cast<const float4 *>(p + i); }
__global__ void __launch_bounds__(256) k_lincomb4(DevGeom g, float *out, Lin L) {
    const int64_t i = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) * 4;
    const int64_t S = 4 * g.nC + g.nF;
    if (i >= S) return;
    const int p = i < 4 * g.nC ? (int)(i / g.nC) : 4;
    float4 x[6];
#pragma unroll
    for (int t = 0; t < 6; ++t) x[t] = t < L.nx ? lin_ld4(L.x[t], i) : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if ((L.planes >> p) & 1u) {
        float4 v = L.acc ? lin_ld4(out, i) : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
#pragma unroll
        for (int t = 0; t < 6; ++t) {
            if (t >= L.nx) break;
            v.x = fmaf(L.a[t], x[t].x, v.x);
            v.y = fmaf(L.a[t], x[t].y, v.y);
            v.z = fmaf(L.a[t], x[t].z, v.z);
            v.w = fmaf(L.a[t], x[t].w, v.w);
        }
        if (p == 4 && L.fa) {
            const int64_t q = i - 4 * g.nC;
            const int64_t f = q / g.ncol;  // the 4 entries share the face row (ncol % 4 == 0)
            if (f >= 1 && f <= g.nlev - 1) {
                const float4 a = lin_ld4(L.fa, q), b = lin_ld4(L.fb, q - g.ncol);
                v.x += a.x + b.x;
                v.y += a.y + b.y;
                v.z += a.z + b.z;
                v.w += a.w + b.w;
            }
        }
        *reinterpret_cast<float4 *>(out + i) = v;
    }
    if (L.out2 && ((L.planes2 >> p) & 1u)) {
        float4 v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
#pragma unroll
        for (int t = 0; t < 6; ++t) {
            if (t >= L.nx) break;
            v.x = fmaf(L.a2[t], x[t].x, v.x);
            v.y = fmaf(L.a2[t], x[t].y, v.y);
            v.z = fmaf(L.a2[t], x[t].z, v.z);
            v.w = fmaf(L.a2[t], x[t].w, v.w);
        }
        *reinterpret_cast<float4 *>(L.out2 + i) = v;
    }
}

// u_3 cotangent of the interior faces += the Laplacian reverse's face terms
__global__ void __launch_bounds__(256) k_fold_faces(DevGeom g, float *Yb, const float *fa, const float *fb) {
    const int col = blockIdx.x * blockDim.x + threadIdx.x;
    const int f = blockIdx.y + 1;
    if (col >= g.ncol || f > g.nlev - 1) return;
    const int64_t q = (int64_t)f * g.ncol + col;
    Yb[4 * g.nC + q] += fa[q] + fb[q - g.ncol];
}

// D0 cotangent per column: sum of the chunk partials
__global__ void k_d0_sum(int ncol, int nchunks, const double *acc, float *out) {
    const int col = blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= ncol) return;
    double s = 0.0;
    for (int c = 0; c < nchunks; ++c) s += acc[(int64_t)c * ncol + col];
    out[col] = (float)s;
}

}  // namespace sadj
