// ca_spectral.hip -- the explicit ("remaining") tendency of the dry dycore:
// horizontal spectral-element dynamics, fourth-order hyperdiffusion,
// vector-invariant vertical advection with deep-atmosphere Coriolis, and the
// dry additional tendencies (Rayleigh u_h sponge, K-theory vertical diffusion).
//
// Restated from (reference paths):
//   src/prognostic_equations/remaining_tendency.jl:48-58
//   src/prognostic_equations/advection.jl:36-91, 205-290
//   src/prognostic_equations/hyperdiffusion.jl:116-147, 247-307
//   src/prognostic_equations/vertical_diffusion_boundary_layer.jl:63-103
//   src/parameterized_tendencies/sponge/rayleigh_sponge.jl:39-44
//
// Mapping: one lane per horizontal node, the 16 nodes of an element in 16
// consecutive lanes (one DPP row); horizontal derivatives are DPP lane moves,
// vertical stencils are a register window carried up the column.
#include "ca_kernels.h"

__device__ __forceinline__ float u3f_at(const DevGeom &g, const float *u3, int f, int col) {
    return (f <= 0 || f >= g.nlev) ? 0.0f : u3[(int64_t)f * g.ncol + col];
}

// ----------------------------------------------------------- prep kernel
// prep_hyperdiffusion_tendency! (hyperdiffusion.jl:141-147, dry):
//   L_u = C123(wgrad(div u)) - C123(wcurl(C123(curl u)))   (3 planes)
//   L_s = wdiv(grad(s_d - s_d,r))                          (1 plane)
// Work item = (node, chunk of levels): the thread walks its chunk upwards,
// carrying u_3 of the face between two levels and the column's horizontal
// metric in registers (each face and metric value is read once); chunk = 0
// means the whole column.
__global__ void __launch_bounds__(256) k_laplacians(DevGeom g, DevNum n, const float *Y, float *L, int chunk) {
    const int col = blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= g.ncol) return;
    const Nod q = ca_node(threadIdx.x & 63);
    const int nl = g.nlev;
    const int64_t nC = g.nC, nc = g.ncol;
    const float *u3 = Y + 4 * nC;
    const int k0 = chunk > 0 ? (int)blockIdx.y * chunk : 0;
    const int k1 = chunk > 0 ? min(k0 + chunk, nl) : nl;
    const float W = c_W[q.j * 4 + q.i], iW = c_iW[q.j * 4 + q.i];
    const float Jh = g.Jh[col], iJh = g.iJh[col];
    const ColMet cm = ca_colmet(g, col);
    const uint32_t cb = (uint32_t)col * 4u;  // the lane's byte offset in every level row
    float u3a = u3f_at(g, u3, k0, col);
    for (int k = k0; k < k1; ++k) {
        const float *Yk = Y + (int64_t)k * nc;
        const float u3b = k + 1 <= nl - 1 ? ca_ldo(u3 + (int64_t)(k + 1) * nc, cb) : 0.0f;
        float o[4];
        lap_point(g, n, q, W, iW, Jh, iJh, cm, k, ca_ldo(Yk, cb), ca_ldo(Yk + nC, cb), ca_ldo(Yk + 2 * nC, cb),
                  ca_ldo(Yk + 3 * nC, cb), u3a, u3b, o);
        float *Lk = L + (int64_t)k * nc;
#pragma unroll
        for (int p = 0; p < 4; ++p) ca_sto(Lk + p * nC, cb, o[p]);
        u3a = u3b;
    }
}

// --------------------------------------------------------- explicit kernel
struct ColState {
    float rhoJ, irhoJ, Kh, Kv, ct1, ct2, u1, u2, rho;  // K = Kh + Kv (horizontal, vertical parts)
    float l4;                  // C3 of nabla^4 u (hyperdiffusion)
    float dr, d1, d2, de;      // partial centre tendencies
    float sd;                  // dry static energy (vertical diffusion)
    float uc0, uc1, uc2;       // velocity in the diffusion basis
    float Kd;                  // eddy diffusivity K_h of the centre
    float rhoe;                // rho e_tot (the final combination's U_4 value)
};

// Work item = (node, chunk of levels): chunk c owns centres and faces
// [c * chunk, (c + 1) * chunk); it recomputes the centre below its first
// level (for the face stencils) and the centre above its last level (for the
// face term of its top centre).  chunk = 0 means the whole column.
// SPH / HYPER / VDIFF select the sphere (horizontal operators), the
// hyperdiffusion and the vertical diffusion at compile time.
// The sphere dynamics with hyperdiffusion and no vertical diffusion (the
// large-grid case) is compiled for 6 wavefronts per SIMD (80 VGPRs, a 20-byte
// spill) instead of the 5 that 93 VGPRs allow: its wavefronts are long and
// few, so how evenly they fill the slots matters more than the spill.  With
// 7-level chunks (the default for large grids) the grid fills 6 per SIMD
// almost exactly twice; at h_elem 30, 63 levels this measured 112.7 us
// against 116.1 us for 5 per SIMD at its best chunk (6).
// FIN (the step's last explicit tendency, which also forms the final
// combination) loads three earlier tendencies at the top of each level so
// that they arrive during the level's arithmetic.  Those 15 values in flight
// need registers: at 4 wavefronts per SIMD (110 VGPRs, no spill) the launch
// measured 171 us at h_elem 30, 63 levels; at 5 (96 VGPRs, a 48-byte spill)
// 224 us; with the loads at the store point instead, 280 us; a separate
// tendency and combination kernel took 112 + 139 us.
template <bool SPH, bool HYPER, bool VDIFF, bool FIN>
__global__ void __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu(SPH && HYPER && !VDIFF ? (FIN ? 4 : 6) : 1, 8)))
k_explicit(DevGeom g, DevNum n, const float *Y, const float *L, float *Yt, int chunk, FinFuse fin) {
    int col = blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= g.ncol) return;
    const Nod q = ca_node(threadIdx.x & 63);
    const int nl = g.nlev;
    const int64_t nC = g.nC, nc = g.ncol;
    const float *u3 = Y + 4 * nC;
    constexpr bool sph = SPH;
    const float W = c_W[q.j * 4 + q.i], iW = c_iW[q.j * 4 + q.i];
    const float Jh = g.Jh[col], iJh = g.iJh[col];
    const float f3n = g.f3n[col], f12a = g.f12a[col], f12b = g.f12b[col];
    const ColMet cm = ca_colmet(g, col);
    const uint32_t cb = (uint32_t)col * 4u;  // the lane's byte offset in every level row
    float spd0 = 0.0f;  // the lowest centre's wind speed (VerticalDiffusion's closure)
    if (VDIFF && n.vdk == 1) {
        const float a1 = ca_ldo(Y + nC, cb), a2 = ca_ldo(Y + 2 * nC, cb);
        float c1, c2;
        ca_contra(cm, g.is2c[0], a1, a2, c1, c2);
        spd0 = sqrtf(a1 * c1 + a2 * c2);
    }
    ColState prev, cur;
    prev.rhoJ = prev.irhoJ = prev.Kh = prev.Kv = prev.ct1 = prev.ct2 = prev.u1 = prev.u2 = prev.rho = prev.l4 = 0.0f;
    prev.dr = prev.d1 = prev.d2 = prev.de = prev.sd = prev.uc0 = prev.uc1 = prev.uc2 = prev.Kd = prev.rhoe = 0.0f;
    float Gp1 = 0.0f, Gp2 = 0.0f;      // G at the face below the current centre
    float Fs_prev = 0.0f, Fu0_prev = 0.0f, Fu1_prev = 0.0f, Fu2_prev = 0.0f;  // diffusive fluxes
    const int k0 = chunk > 0 ? (int)blockIdx.y * chunk : 0;
    const int k1 = chunk > 0 ? min(k0 + chunk, nl) : nl;
    const int kbeg = k0 > 0 ? k0 - 1 : 0;
    const int kend = k1;  // inclusive: face k1 closes the top owned centre
    float u3a = u3f_at(g, u3, kbeg, col);  // u_3 at face k
    for (int k = kbeg; k <= kend; ++k) {
        // FIN: the earlier stages' tendencies at centre k - 1 and face k, loaded
        // here so that their latency hides behind this level's arithmetic
        float fe_c[3][4], fe_f[3];
        if (FIN) {
            if (k - 1 >= k0 && k - 1 < k1)
#pragma unroll
                for (int j = 0; j < 3; ++j)
#pragma unroll
                    for (int c = 0; c < 4; ++c) fe_c[j][c] = ca_ldo(fin.fe[j] + c * nC + (int64_t)(k - 1) * nc, cb);
            if ((k >= k0 && k < k1) || (k == nl && k1 == nl))
#pragma unroll
                for (int j = 0; j < 3; ++j) fe_f[j] = ca_ldo(fin.fe[j] + 4 * nC + (int64_t)k * nc, cb);
        }
        float u3b = (k + 1 >= 1 && k + 1 <= nl - 1) ? ca_ldo(u3 + (int64_t)(k + 1) * nc, cb) : 0.0f;
        if (k < nl) {
            const float *Yk = Y + (int64_t)k * nc;
            float rho = ca_ldo(Yk, cb), u1 = ca_ldo(Yk + nC, cb), u2 = ca_ldo(Yk + 2 * nC, cb),
                  rhoe = ca_ldo(Yk + 3 * nC, cb);
            float s = g.sc[k], s2 = s * s, is2 = g.is2c[k], Jv = g.Jvc[k];
            const Met m = ca_met(W, iW, Jh * s2 * Jv, iJh * g.isJc[k]);
            const float J = m.J;
            float ct1, ct2;
            ca_contra(cm, is2, u1, u2, ct1, ct2);
            float ga = g.g33f[k], gb = g.g33f[k + 1];
            const float Khor = 0.5f * (u1 * ct1 + u2 * ct2);
            const float Kver = 0.25f * (u3a * ga * u3a + u3b * gb * u3b);
            float K = Khor + Kver;
            // temperature from e_int in float64 (e_tot and Phi are large)
            const double e_tot_d = (double)rhoe * ca_drcp((double)rho);
            const double Td =
                fmax(n.dT_0 + (e_tot_d - (double)K - g.Phic_d[k] + n.dR_d * n.dT_0) * n.dicv_d, n.dTmin);
            float T = (float)Td;
            float h = (float)(e_tot_d + n.dR_d * Td);
            float p = (float)((double)rho * n.dR_d * Td);
            float Pi = ca_exner(n, p);
            float dr = 0.0f, d1 = 0.0f, d2 = 0.0f, de = 0.0f, l4 = 0.0f;
            // The chunk's halo centres (k0 - 1 below, k1 above; uniform per
            // workgroup) only feed the face stencils of faces k0 and k1: they
            // need K, the velocities, rho J, l4 and the diffusion inputs, not
            // their own tendencies.
            const bool own = k >= k0 && k < k1;
            if (sph && own) {
                // ---- horizontal_dynamics_tendency! (advection.jl:36-91)
                float m1 = rho * ct1, m2 = rho * ct2;
                float wd = op_wdiv(q, m, m1, m2);
                dr -= 0.5f * wd + 0.5f * wd;  // split_div(rho u, 1)
                // split_div(rho u, h) = 1/2 wdiv(rho u h) + 1/2 (h wdiv(rho u) + rho u . grad h),
                // with h taken relative to the element's first node (exact: linear in constants)
                float h0 = ca_elem0(h);
                float hh = h - h0;
                float wdh = op_wdiv(q, m, m1 * hh, m2 * hh);
                float gh1, gh2;
                op_grad(q, hh, gh1, gh2);
                de -= 0.5f * wdh + 0.5f * (hh * wd + (m1 * gh1 + m2 * gh2)) + h0 * wd;
                // u_h -= C12(grad(K + Phi - Phi_r) + cp_d (th' grad Pi + grad(th' Pi) - Pi grad th') / 2)
                float p0 = ca_elem0(p), Pi0 = ca_elem0(Pi), ip0 = ca_rcp(p0);
                float dPhr, dPi;
                ca_dref(n, p0, ip0, Pi0, p, dPhr, dPi);
                float A = K - dPhr;
                float thPi = T - ca_Tref(n, Pi);
                float th = thPi * ca_rcp(Pi);
                float g1, g2, a1, a2, b1, b2, c1, c2;
                op_grad(q, A, g1, g2);
                op_grad(q, dPi, a1, a2);
                op_grad(q, thPi, b1, b2);
                op_grad(q, th, c1, c2);
                d1 -= g1 + n.cp_d * (th * a1 + b1 - Pi * c1) * 0.5f;
                d2 -= g2 + n.cp_d * (th * a2 + b2 - Pi * c2) * 0.5f;
            }
            if (sph && HYPER && !own) {  // l4 alone, as below
                const float X3 = ca_ldo(L + 2 * nC + (int64_t)k * nc, cb);
                float cr1, cr2;
                op_curl12(q, m, X3, cr1, cr2);
                float cc1, cc2;
                ca_cov(cm, s2, cr1, cr2, cc1, cc2);
                const float w3 = op_wcurl3(q, m, cc1, cc2);
                const float wc3 = Jv * Jv * w3;
                l4 = n.dd * 0.0f - wc3;
            }
            if (sph && own) {
                // ---- apply_hyperdiffusion_tendency! (hyperdiffusion.jl:273-307)
                if (HYPER) {
                    const float *Lk = L + (int64_t)k * nc;
                    float X1 = ca_ldo(Lk, cb), X2 = ca_ldo(Lk + nC, cb), X3 = ca_ldo(Lk + 2 * nC, cb),
                          Ls = ca_ldo(Lk + 3 * nC, cb);
                    float v1, v2;
                    ca_contra(cm, is2, X1, X2, v1, v2);
                    float dv = op_div(q, m, v1, v2);
                    float gd1, gd2;
                    op_wgrad(q, m, dv, gd1, gd2);
                    float cr3 = op_curl3(q, m, X1, X2);
                    float cr1, cr2;
                    op_curl12(q, m, X3, cr1, cr2);
                    float cc1, cc2;
                    ca_cov(cm, s2, cr1, cr2, cc1, cc2);
                    float cc3 = Jv * Jv * cr3;
                    float w3 = op_wcurl3(q, m, cc1, cc2);
                    float w1, w2;
                    op_wcurl12(q, m, cc3, w1, w2);
                    float wc1, wc2;
                    ca_cov(cm, s2, w1, w2, wc1, wc2);
                    float wc3 = Jv * Jv * w3;
                    d1 -= n.nu_v * (n.dd * gd1 - wc1);
                    d2 -= n.nu_v * (n.dd * gd2 - wc2);
                    l4 = n.dd * 0.0f - wc3;
                    float s1, s2g;
                    op_grad(q, Ls, s1, s2g);
                    float r1, r2;
                    ca_contra(cm, is2, rho * s1, rho * s2g, r1, r2);
                    de -= n.nu_s * op_wdiv(q, m, r1, r2);
                }
            }
            if (own) {
                // ---- explicit_vertical_advection_tendency! centre part (advection.jl:273-277)
                float om3 = sph ? op_wcurl3(q, m, u1, u2) : 0.0f;
                float a3 = f3n * g.iJvc[k] + om3;
                d1 -= J * (0.0f - a3 * ct2);
                d2 -= J * (a3 * ct1 - 0.0f);
                // ---- Rayleigh sponge on u_h (rayleigh_sponge.jl:39-44)
                if (n.rayleigh && n.alpha_uh != 0.0f) {
                    float beta = n.alpha_uh * g.zeta_c[k];
                    d1 += -beta * u1;
                    d2 += -beta * u2;
                }
            }
            cur.rhoJ = rho * J;
            cur.irhoJ = ca_rcp(cur.rhoJ);
            cur.Kh = Khor;
            cur.Kv = Kver;
            cur.ct1 = ct1;
            cur.ct2 = ct2;
            cur.u1 = u1;
            cur.u2 = u2;
            cur.rho = rho;
            if (FIN) cur.rhoe = rhoe;
            cur.l4 = l4;
            cur.dr = dr;
            cur.d1 = d1;
            cur.d2 = d2;
            cur.de = de;
            if (VDIFF) {
                cur.sd = n.cp_d * (T - n.T_0) + g.Phic[k];
                cur.Kd = ca_Kc(g, n, col, k, p, spd0);
                if (sph) {
                    const float *B = g.basis;
                    float is = ca_rcp(s);
                    cur.uc0 = (u1 * B[0 * nc + col] + u2 * B[3 * nc + col]) * is;
                    cur.uc1 = (u1 * B[1 * nc + col] + u2 * B[4 * nc + col]) * is;
                    cur.uc2 = (u1 * B[2 * nc + col] + u2 * B[5 * nc + col]) * is;
                } else {
                    cur.uc0 = u1;
                    cur.uc1 = u2;
                    cur.uc2 = 0.0f;
                }
            }
        }
        // ---- face k (advection.jl:278): u_3 tendency and the face term G
        float G1 = 0.0f, G2 = 0.0f, du3 = 0.0f;
        float Fs = 0.0f, Fu0 = 0.0f, Fu1 = 0.0f, Fu2 = 0.0f;
        {
            // the horizontal weak curl of u_3 at face k needs every lane of the element
            float x = 0.0f, y = 0.0f;
            const bool interior = (k >= 1 && k <= nl - 1);
            const float sf = g.sf[k];
            const Met mf = ca_met(W, iW, Jh * sf * sf * g.Jvf[k], iJh * g.isJf[k]);
            const float Jf = mf.J, iJf = mf.iJ;
            if (sph && interior) op_wcurl12(q, mf, u3a, x, y);
            if (interior) {
                float fu3 = u3a * g.g33f[k];
                float wa = -(cur.u2 - prev.u2) * iJf + x;
                float wb = (cur.u1 - prev.u1) * iJf + y;
                const float isf = g.isf[k];
                float A1 = f12a * isf + wa, A2 = f12b * isf + wb;
                float rJf = 0.5f * (prev.rhoJ + cur.rhoJ);
                float b3 = rJf * fu3;
                G1 = Jf * (A2 * b3);
                G2 = Jf * (0.0f - A1 * b3);
                float i1 = 0.5f * (prev.ct1 + cur.ct1), i2 = 0.5f * (prev.ct2 + cur.ct2);
                // grad K differenced part by part: the vertical part is a small
                // difference riding on the horizontal kinetic energy
                du3 = -(Jf * (A1 * i2 - A2 * i1) + ((cur.Kh - prev.Kh) + (cur.Kv - prev.Kv)));
                if (SPH && HYPER) {
                    float wA = prev.rhoJ, wB = cur.rhoJ;
                    du3 -= n.nu_v * ((0.5f * (wA * prev.l4 + wB * cur.l4)) * ca_rcp(0.5f * (wA + wB)));
                }
                if (VDIFF) {
                    // vertical_diffusion_boundary_layer.jl:86-103
                    float rK = ca_rhoK(prev.rho, cur.rho, prev.Kd, cur.Kd);
                    float iJv = g.iJvf[k];
                    Fs = Jf * (iJv * iJv * (-(rK * (cur.sd - prev.sd))));
                    float fac = Jf * iJv * (-2.0f * rK) * 0.5f;
                    Fu0 = fac * (cur.uc0 - prev.uc0) * iJv;
                    Fu1 = fac * (cur.uc1 - prev.uc1) * iJv;
                    Fu2 = fac * (cur.uc2 - prev.uc2) * iJv;
                }
            }
            if ((k >= k0 && k < k1) || (k == nl && k1 == nl)) {
                float v = du3;
                if (FIN) {  // U_4 + dt sum_j c_j F_E,j at face k (u3a: U_4's u_3 there)
                    float acc = fin.cg * du3;
                    acc = fmaf(fin.c[0], fe_f[0], acc);
                    acc = fmaf(fin.c[1], fe_f[1], acc);
                    acc = fmaf(fin.c[2], fe_f[2], acc);
                    v = u3a + acc;
                }
                ca_sto(Yt + 4 * nC + (int64_t)k * nc, cb, v);
            }
        }
        // ---- centre k-1 is complete
        if (k - 1 >= k0 && k - 1 < k1) {
            const int kk = k - 1;
            float d1 = prev.d1 - (0.5f * (Gp1 + G1)) * prev.irhoJ;
            float d2 = prev.d2 - (0.5f * (Gp2 + G2)) * prev.irhoJ;
            float de = prev.de;
            if (VDIFF) {
                float s = g.sc[kk];
                const float iJc = iJh * g.isJc[kk];
                de -= (Fs - Fs_prev) * iJc;
                float e0 = (Fu0 - Fu0_prev) * iJc, e1 = (Fu1 - Fu1_prev) * iJc, e2 = (Fu2 - Fu2_prev) * iJc;
                float c1, c2;
                if (sph) {
                    const float *B = g.basis;
                    c1 = s * (e0 * B[6 * nc + col] + e1 * B[7 * nc + col] + e2 * B[8 * nc + col]);
                    c2 = s * (e0 * B[9 * nc + col] + e1 * B[10 * nc + col] + e2 * B[11 * nc + col]);
                } else {
                    c1 = e0;
                    c2 = e1;
                }
                const float irho = ca_rcp(prev.rho);
                d1 -= c1 * irho;
                d2 -= c2 * irho;
            }
            float *Ytk = Yt + (int64_t)kk * nc;
            if (FIN) {  // U_4 + dt sum_j c_j F_E,j at centre kk
                const float v[4] = {prev.dr, d1, d2, de}, u[4] = {prev.rho, prev.u1, prev.u2, prev.rhoe};
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    float acc = fin.cg * v[c];
                    acc = fmaf(fin.c[0], fe_c[0][c], acc);
                    acc = fmaf(fin.c[1], fe_c[1][c], acc);
                    acc = fmaf(fin.c[2], fe_c[2][c], acc);
                    ca_sto(Ytk + c * nC, cb, u[c] + acc);
                }
            } else {
                ca_sto(Ytk, cb, prev.dr);
                ca_sto(Ytk + nC, cb, d1);
                ca_sto(Ytk + 2 * nC, cb, d2);
                ca_sto(Ytk + 3 * nC, cb, de);
            }
        }
        Gp1 = G1;
        Gp2 = G2;
        Fs_prev = Fs;
        Fu0_prev = Fu0;
        Fu1_prev = Fu1;
        Fu2_prev = Fu2;
        prev = cur;
        u3a = u3b;
    }
}

template __global__ void k_explicit<true, true, false, false>(DevGeom, DevNum, const float *, const float *,
                                                                    float *, int, FinFuse);
template __global__ void k_explicit<true, true, false, true>(DevGeom, DevNum, const float *, const float *,
                                                                    float *, int, FinFuse);
template __global__ void k_explicit<true, false, false, false>(DevGeom, DevNum, const float *, const float *,
                                                                    float *, int, FinFuse);
template __global__ void k_explicit<true, false, false, true>(DevGeom, DevNum, const float *, const float *,
                                                                    float *, int, FinFuse);
template __global__ void k_explicit<true, true, true, false>(DevGeom, DevNum, const float *, const float *,
                                                                    float *, int, FinFuse);
template __global__ void k_explicit<true, true, true, true>(DevGeom, DevNum, const float *, const float *,
                                                                    float *, int, FinFuse);
template __global__ void k_explicit<true, false, true, false>(DevGeom, DevNum, const float *, const float *,
                                                                    float *, int, FinFuse);
template __global__ void k_explicit<true, false, true, true>(DevGeom, DevNum, const float *, const float *,
                                                                    float *, int, FinFuse);
template __global__ void k_explicit<false, false, false, false>(DevGeom, DevNum, const float *, const float *,
                                                                    float *, int, FinFuse);
template __global__ void k_explicit<false, false, false, true>(DevGeom, DevNum, const float *, const float *,
                                                                    float *, int, FinFuse);
template __global__ void k_explicit<false, false, true, false>(DevGeom, DevNum, const float *, const float *,
                                                                    float *, int, FinFuse);
template __global__ void k_explicit<false, false, true, true>(DevGeom, DevNum, const float *, const float *,
                                                                    float *, int, FinFuse);

// ------------------------------------------------------------ surface flux
// surface_flux_tendency! (surface_flux.jl:101-145) with the surface conditions
// of update_surface_conditions! (surface_conditions.jl:22-70, 200-288) for
// ExchangeCoefficients(Cd, Ch), dry air and no boundary overrides, added to
// the explicit tendency Yt at the first centre of every column (a thread per
// column, after k_explicit and before the tendency's DSS).  A kernel of its
// own: folded into k_explicit's level loop it raised that kernel's spill from
// 20 to 60 bytes per lane even with the fluxes off.
//   interior wind   CT1 / CT2 components times |a_1|, |a_2| (projected_vector_data)
//   U               max(|u|, gustiness)
//   rho_s           the interior pressure extrapolated hydrostatically over
//                   dz at the mean of the two temperatures, over R_d T_s
//   shf             -rho_s Ch U (cp_d (T - T_s) + g dz)
//   momentum flux   -rho_s Cd U (s_f0 / s_c0) u_i in covariant components (built
//                   from unit CT1 / CT2 vectors at the surface, lowered with the
//                   surface metric)
// Both enter through the bottom face of a divergence with zero interior flux:
// the tendency is the flux times J_f(surface) / J_c = s_f0^2 / (s_c0^2 dz_c0).
// scale multiplies what is added: 1 for the tendency, dt b_4 when Yt holds
// the step's final combination (k_explicit with FIN).
__global__ void __launch_bounds__(256) k_sfc_flux(DevGeom g, DevNum n, const float *Y, float *Yt, float scale) {
    const int col = blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= g.ncol) return;
    const CV c = cv_load(g, n, Y, 0, col);
    const float s = g.sc[0], s0 = g.sf[0];
    const float ui = c.ct1 * sqrtf(g.gd11[col]) * s, vi = c.ct2 * sqrtf(g.gd22[col]) * s;
    const float U = fmaxf(sqrtf(ui * ui + vi * vi), n.gust);
    const float T = c.T, Ts = g.Tsfc[col];
    const float dz = g.zc[0] - g.zf[0];
    const float rho_s = c.rho * (T / Ts) * expf(n.grav * dz / (n.R_d * (0.5f * (T + Ts))));
    const float shf = -rho_s * n.Ch * U * (n.cp_d * (T - Ts) + n.grav * dz);
    const float fac = scale * s0 * s0 * g.isJc[0];
    const float cD = fac * rho_s * n.Cd * U * (s0 / s) / c.rho;
    const int64_t nC = g.nC;
    Yt[nC + col] -= cD * c.u1;
    Yt[2 * nC + col] -= cD * c.u2;
    Yt[3 * nC + col] += fac * shf;
}
