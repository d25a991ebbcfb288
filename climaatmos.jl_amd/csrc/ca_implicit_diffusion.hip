// ca_implicit_diffusion.hip -- vertical diffusion treated implicitly
// (diff_mode Implicit, YAML implicit_diffusion: true) for the dry
// DecayWithHeightDiffusion closure (K_u = K_h = D0 per column, H = Inf).
//
//   tendency   vertical_diffusion_boundary_layer_tendency! moves from
//              additional_tendency! to implicit_tendency! (implicit_tendency.jl:73-82;
//              vertical_diffusion_boundary_layer.jl:63-103)
//   blocks     update_diffusion_jacobian! (manual_sparse_jacobian.jl:1031-1261):
//              D = advdiv_matrix . diag(I^f(rho) / I^f(1 / max(K, eps))) . grad_matrix,
//              (rho e_tot, rho e_tot) = dtg D diag(cp_d / (cv_d rho)) - I,
//              (u_h, u_h) = dtg diag(1 / rho) D - I, (rho e_tot, rho) = 0
//   solve      ApproximateBlockArrowheadIterativeSolve (manual_sparse_jacobian.jl:538-578):
//              a Schur complement reduction on the scalars (rho, rho e_tot), whose
//              block is inverted exactly, and n_iters stationary iterations from
//              zero on the velocity Schur complement, preconditioned by the
//              complement formed with the scalar block's main diagonal and solved
//              u_h first, then u_3 (ClimaCore MatrixFields, restated).
//
// A column per thread: the tridiagonal solves are Thomas sweeps over
// column-strided scratch planes (coalesced across the threads of a
// wavefront).  The fused step takes an implicit stage as a sequence of these
// kernels (assembly, implicit tendency, Jacobian, solve, update, upwind
// correction, F_I), and F_I then carries u_h as well.
#include "ca_kernels.h"

// vertical diffusion fluxes through interior face f of one column: the
// s_d flux (J_f g^33-weighted) and the Cartesian momentum fluxes, as the
// explicit tendency forms them (k_explicit, VDIFF)
__device__ __forceinline__ void vdiff_face(const DevGeom &g, const DevNum &n, int col, int f, const CV &a, const CV &b,
                                           float Ka, float Kb, float &Fs, float &F0, float &F1, float &F2) {
    const int64_t nc = g.ncol;
    const float rK = ca_rhoK(a.rho, b.rho, Ka, Kb);
    const float sf = g.sf[f];
    const float Jf = g.Jh[col] * sf * sf * g.Jvf[f];
    const float iJv = g.iJvf[f];
    // s_d(b) - s_d(a) = cp_d (T_b - T_a) + (Phi_b - Phi_a), formed from the
    // float64 temperatures and geopotential step: s_d itself is ~3e5 J/kg and
    // its level differences ~1e3, which float32 values of s_d would lose
    const float dsd = n.cp_d * (float)(b.Td - a.Td) + (float)g.dPhif_d[f];
    Fs = Jf * (iJv * iJv * (-(rK * dsd)));
    // Cartesian velocity differences (u_h / s in the Cartesian basis) formed
    // from the covariant differences plus the metric step, not as a
    // difference of two nearly equal Cartesian values
    float du[3];
    if (n.sphere) {
        const float *B = g.basis;
        const float isa = 1.0f / g.sc[f - 1], isb = 1.0f / g.sc[f];
        const float dis = (g.sc[f - 1] - g.sc[f]) * isa * isb;  // 1/s_b - 1/s_a
        const float d1 = b.u1 - a.u1, d2 = b.u2 - a.u2;
        for (int d = 0; d < 3; ++d) {
            const float B1 = B[d * nc + col], B2 = B[(3 + d) * nc + col];
            du[d] = (d1 * B1 + d2 * B2) * isb + (a.u1 * B1 + a.u2 * B2) * dis;
        }
    } else {
        du[0] = b.u1 - a.u1;
        du[1] = b.u2 - a.u2;
        du[2] = 0.0f;
    }
    const float fac = Jf * iJv * (-2.0f * rK) * 0.5f;
    F0 = fac * du[0] * iJv;
    F1 = fac * du[1] * iJv;
    F2 = fac * du[2] * iJv;
}

// the lowest centre's horizontal wind speed |u_h| = sqrt(u_i u^i) (VerticalDiffusion's
// closure; 0 for DecayWithHeightDiffusion, which does not read it)
__device__ __forceinline__ float spd_lowest(const DevGeom &g, const DevNum &n, const float *Y, int col) {
    if (n.vdk != 1) return 0.0f;
    const float u1 = Y[g.nC + col], u2 = Y[2 * g.nC + col];
    float c1, c2;
    ca_contra(g, col, g.is2c[0], u1, u2, c1, c2);
    return sqrtf(u1 * c1 + u2 * c2);
}

// Yt += the vertical diffusion tendency (rho e_tot, u_h) at centre l of every column
__global__ void __launch_bounds__(128) k_vdiff_add(DevGeom g, DevNum n, const float *Y, float *Yt) {
    const int col = blockIdx.x * blockDim.x + threadIdx.x, l = blockIdx.y;
    if (col >= g.ncol || l >= g.nlev) return;
    const int nl = g.nlev;
    const int64_t nC = g.nC, nc = g.ncol;
    const CV c = cv_load(g, n, Y, l, col);
    const float spd0 = spd_lowest(g, n, Y, col);
    const float Kc = ca_Kc(g, n, col, l, c.p, spd0);
    float Fs[2] = {0.0f, 0.0f}, Fu[2][3] = {{0.0f, 0.0f, 0.0f}, {0.0f, 0.0f, 0.0f}};
    if (l >= 1) {
        const CV a = cv_load(g, n, Y, l - 1, col);
        vdiff_face(g, n, col, l, a, c, ca_Kc(g, n, col, l - 1, a.p, spd0), Kc, Fs[0], Fu[0][0], Fu[0][1], Fu[0][2]);
    }
    if (l + 1 <= nl - 1) {
        const CV b = cv_load(g, n, Y, l + 1, col);
        vdiff_face(g, n, col, l + 1, c, b, Kc, ca_Kc(g, n, col, l + 1, b.p, spd0), Fs[1], Fu[1][0], Fu[1][1],
                   Fu[1][2]);
    }
    const float iJc = g.iJh[col] * g.isJc[l];
    const int64_t q = (int64_t)l * nc + col;
    Yt[3 * nC + q] -= (Fs[1] - Fs[0]) * iJc;
    const float e0 = (Fu[1][0] - Fu[0][0]) * iJc, e1 = (Fu[1][1] - Fu[0][1]) * iJc, e2 = (Fu[1][2] - Fu[0][2]) * iJc;
    float c1, c2;
    if (n.sphere) {
        const float *B = g.basis;
        const float s = g.sc[l];
        c1 = s * (e0 * B[6 * nc + col] + e1 * B[7 * nc + col] + e2 * B[8 * nc + col]);
        c2 = s * (e0 * B[9 * nc + col] + e1 * B[10 * nc + col] + e2 * B[11 * nc + col]);
    } else {
        c1 = e0;
        c2 = e1;
    }
    Yt[nC + q] -= c1 / c.rho;
    Yt[2 * nC + q] -= c2 / c.rho;
}

// the diffusion blocks at centre row l: Aee and Ahh (lo, d, up)
__global__ void __launch_bounds__(128) k_diffusion_jacobian(DevGeom g, DevNum n, const float *Y, float dtg,
                                                            DiffBands B) {
    const int col = blockIdx.x * blockDim.x + threadIdx.x, l = blockIdx.y;
    if (col >= g.ncol || l >= g.nlev) return;
    const int nl = g.nlev;
    const int64_t nc = g.ncol;
    const float *rho = Y;
    const int64_t q = (int64_t)l * nc + col;
    const float r0 = rho[q];
    const float Jc = g.Jh[col] * g.sc[l] * g.sc[l] * g.Jvc[l];
    // centre diffusivities (VerticalDiffusion's depend on the pressure)
    const float spd0 = spd_lowest(g, n, Y, col);
    auto Kat = [&](int k) { return ca_Kc(g, n, col, k, n.vdk == 1 ? cv_load(g, n, Y, k, col).p : 0.0f, spd0); };
    const float K0 = Kat(l);
    float dlo = 0.0f, dup = 0.0f, rm = 1.0f, rp = 1.0f;
    if (l >= 1) {
        rm = rho[q - nc];
        const float sf = g.sf[l];
        dlo = (g.Jh[col] * sf * sf * g.Jvf[l]) * g.g33f[l] * ca_rhoK(rm, r0, Kat(l - 1), K0) / Jc;
    }
    if (l + 1 <= nl - 1) {
        rp = rho[q + nc];
        const float sf = g.sf[l + 1];
        dup = (g.Jh[col] * sf * sf * g.Jvf[l + 1]) * g.g33f[l + 1] * ca_rhoK(r0, rp, K0, Kat(l + 1)) / Jc;
    }
    const float dd = -(dlo + dup);
    const float kc = n.cp_d / n.cv_d;
    B.elo[q] = l >= 1 ? dtg * dlo * (kc / rm) : 0.0f;
    B.ed[q] = dtg * dd * (kc / r0) - 1.0f;
    B.eup[q] = l + 1 <= nl - 1 ? dtg * dup * (kc / rp) : 0.0f;
    B.hlo[q] = dtg * (1.0f / r0) * dlo;
    B.hd[q] = dtg * (1.0f / r0) * dd - 1.0f;
    B.hup[q] = dtg * (1.0f / r0) * dup;
}

// Thomas sweep over n rows of column-strided planes (element k at p[k * s]):
// lo[0] and up[n - 1] are not read; w is scratch
__device__ __forceinline__ void tri_col(int nr, int64_t s, const float *lo, const float *d, const float *up,
                                        const float *r, float *x, float *w) {
    float m = d[0];
    x[0] = r[0] / m;
    for (int i = 1; i < nr; ++i) {
        const float wi = up[(int64_t)(i - 1) * s] / m;
        w[(int64_t)i * s] = wi;
        m = d[(int64_t)i * s] - lo[(int64_t)i * s] * wi;
        x[(int64_t)i * s] = (r[(int64_t)i * s] - lo[(int64_t)i * s] * x[(int64_t)(i - 1) * s]) / m;
    }
    for (int i = nr - 2; i >= 0; --i) x[(int64_t)i * s] -= w[(int64_t)(i + 1) * s] * x[(int64_t)(i + 1) * s];
}

// ldiv! with implicit diffusion, a column per thread: dY = solve(A, rscale R).
// W: 9 centre planes (ye, xh1, xh2, rh1, rh2, zh1, zh2, g, m) and 7 face
// planes (xw, rw, zw, slo, sd, sup, wk).
__global__ void __launch_bounds__(128) k_ldiv_approx(DevGeom g, Bands A, DiffBands E, int iters, float rscale,
                                                     float *dY, const float *R, float *W) {
    const int col = blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= g.ncol) return;
    const int nl = g.nlev;
    const int64_t nC = g.nC, nF = g.nF, s = g.ncol;
    float *ye = W + col, *xh1 = ye + nC, *xh2 = xh1 + nC, *rh1 = xh2 + nC, *rh2 = rh1 + nC, *zh1 = rh2 + nC,
          *zh2 = zh1 + nC, *gg = zh2 + nC, *mm = gg + nC;
    float *xw = W + 9 * nC + col, *rw = xw + nF, *zw = rw + nF, *slo = zw + nF, *sd = slo + nF, *sup = sd + nF,
          *wk = sup + nF;
    const float *Rr = R + col, *R1 = R + nC + col, *R2 = R + 2 * nC + col, *Re = R + 3 * nC + col,
                *Rw = R + 4 * nC + col;
    auto at = [&](const float *p, int k) { return p[(int64_t)k * s + col]; };
    // y1 = inv(A11) b1: y_rho = -b_rho, A_ee y_e = b_e
    for (int k = 0; k < nl; ++k) gg[(int64_t)k * s] = rscale * Re[(int64_t)k * s];
    tri_col(nl, s, E.elo + col, E.ed + col, E.eup + col, gg, ye, wk);
    // b2' (u_3 rows; kept in rw as the iteration's right-hand side base) and P_ww
    // b2' of the u_3 rows (in dY's u_3 plane, which is written last) and the
    // preconditioner's u_3 block P_ww = A_ww + A_wrho A_rhow - A_we diag(A_ee)^-1 A_ew
    float *bw = dY + 4 * nC + col;
    for (int f = 0; f <= nl; ++f) {
        const int64_t qf = (int64_t)f * s;
        float b = rscale * Rw[qf];
        float lo = at(A.lo, f), d = at(A.d, f), up = at(A.up, f);
        if (f >= 1 && f <= nl - 1) {
            const int a = f - 1, bb = f;
            b -= at(A.Aur0, f) * (-rscale * Rr[(int64_t)a * s]) + at(A.Aur1, f) * (-rscale * Rr[(int64_t)bb * s]) +
                 at(A.Aue0, f) * ye[(int64_t)a * s] + at(A.Aue1, f) * ye[(int64_t)bb * s];
            const float eda = at(E.ed, a), edb = at(E.ed, bb);
            lo += at(A.Aur0, f) * at(A.Aru0, a) - at(A.Aue0, f) * at(A.Aeu0, a) / eda;
            d += at(A.Aur0, f) * at(A.Aru1, a) + at(A.Aur1, f) * at(A.Aru0, bb) -
                 (at(A.Aue0, f) * at(A.Aeu1, a) / eda + at(A.Aue1, f) * at(A.Aeu0, bb) / edb);
            up += at(A.Aur1, f) * at(A.Aru1, bb) - at(A.Aue1, f) * at(A.Aeu1, bb) / edb;
        }
        bw[qf] = b;
        slo[qf] = lo;
        sd[qf] = d;
        sup[qf] = up;
        xw[qf] = 0.0f;
    }
    for (int k = 0; k < nl; ++k) xh1[(int64_t)k * s] = xh2[(int64_t)k * s] = 0.0f;
    for (int it = 0; it < iters; ++it) {
        // r = b2' - S x; S x = A22 x + A_wrho A_rhow x_w - A_we inv(A_ee) A_ew x_w
        for (int k = 0; k < nl; ++k) {
            const int64_t qk = (int64_t)k * s;
            float s1 = at(E.hd, k) * xh1[qk], s2 = at(E.hd, k) * xh2[qk];
            if (k >= 1) {
                s1 += at(E.hlo, k) * xh1[qk - s];
                s2 += at(E.hlo, k) * xh2[qk - s];
            }
            if (k + 1 < nl) {
                s1 += at(E.hup, k) * xh1[qk + s];
                s2 += at(E.hup, k) * xh2[qk + s];
            }
            rh1[qk] = rscale * R1[qk] - s1;
            rh2[qk] = rscale * R2[qk] - s2;
            gg[qk] = at(A.Aeu0, k) * xw[qk] + at(A.Aeu1, k) * xw[qk + s];
        }
        tri_col(nl, s, E.elo + col, E.ed + col, E.eup + col, gg, mm, wk);
        for (int f = 0; f <= nl; ++f) {
            const int64_t qf = (int64_t)f * s;
            float sx = at(A.d, f) * xw[qf];
            if (f >= 1) sx += at(A.lo, f) * xw[qf - s];
            if (f + 1 <= nl) sx += at(A.up, f) * xw[qf + s];
            if (f >= 1 && f <= nl - 1) {
                const int64_t qa = qf - s, qb = qf;
                sx += at(A.Auh0a, f) * xh1[qa] + at(A.Auh0b, f) * xh2[qa] + at(A.Auh1a, f) * xh1[qb] +
                      at(A.Auh1b, f) * xh2[qb];
                const float pa = at(A.Aru0, f - 1) * xw[qa] + at(A.Aru1, f - 1) * xw[qa + s];
                const float pb = at(A.Aru0, f) * xw[qb] + at(A.Aru1, f) * xw[qb + s];
                sx += at(A.Aur0, f) * pa + at(A.Aur1, f) * pb;
                sx -= at(A.Aue0, f) * mm[qa] + at(A.Aue1, f) * mm[qb];
            }
            rw[qf] = bw[qf] - sx;
        }
        // z = inv(P) r: u_h, then u_3 with the (u_3, u_h) coupling
        tri_col(nl, s, E.hlo + col, E.hd + col, E.hup + col, rh1, zh1, wk);
        tri_col(nl, s, E.hlo + col, E.hd + col, E.hup + col, rh2, zh2, wk);
        for (int f = 1; f < nl; ++f) {
            const int64_t qf = (int64_t)f * s;
            rw[qf] -= at(A.Auh0a, f) * zh1[qf - s] + at(A.Auh0b, f) * zh2[qf - s] + at(A.Auh1a, f) * zh1[qf] +
                      at(A.Auh1b, f) * zh2[qf];
        }
        tri_col(nl + 1, s, slo, sd, sup, rw, zw, wk);
        for (int k = 0; k < nl; ++k) {
            xh1[(int64_t)k * s] += zh1[(int64_t)k * s];
            xh2[(int64_t)k * s] += zh2[(int64_t)k * s];
        }
        for (int f = 0; f <= nl; ++f) xw[(int64_t)f * s] += zw[(int64_t)f * s];
    }
    // x1 = inv(A11) (b1 - A12 x2)
    for (int k = 0; k < nl; ++k) {
        const int64_t qk = (int64_t)k * s;
        gg[qk] = rscale * Re[qk] - (at(A.Aeu0, k) * xw[qk] + at(A.Aeu1, k) * xw[qk + s]);
    }
    tri_col(nl, s, E.elo + col, E.ed + col, E.eup + col, gg, mm, wk);
    for (int k = 0; k < nl; ++k) {
        const int64_t qk = (int64_t)k * s;
        dY[qk + col] = at(A.Aru0, k) * xw[qk] + at(A.Aru1, k) * xw[qk + s] - rscale * Rr[qk];
        dY[nC + qk + col] = xh1[qk];
        dY[2 * nC + qk + col] = xh2[qk];
        dY[3 * nC + qk + col] = mm[qk];
    }
    for (int f = 0; f <= nl; ++f) dY[4 * nC + (int64_t)f * s + col] = xw[(int64_t)f * s];
}

// ---------------------------------------------------------------------------
// The whole implicit stage with implicit diffusion as one kernel, built like
// k_newton_pcr: a workgroup assembles a tile of 16 columns into LDS, each
// wavefront then takes one column with lane k = centre k and face k, and the
// results leave through LDS.  Per column: U~, T_imp (advection, u_3 pressure
// gradient and buoyancy, Rayleigh, and the vertical diffusion of rho e_tot and
// u_h), the manual Jacobian and the diffusion blocks, the approximate
// arrowhead solve, the Newton update, the upwind correction and
// F_I = (U - U~) / dtg, all in registers.  The tridiagonal solves are parallel
// cyclic reductions across the 64 lanes: float32 for the diffusion blocks
// A_ee and A_hh (-I minus a diffusion operator: diagonally dominant), float64
// for the u_3 Schur complement (stiff, as in the dry stage).  Same algebra as
// the hook kernels above; the order of the roundings differs.

namespace idp {

// 1 / d from the hardware reciprocal (1 ulp) and one Newton-Raphson step
// r' = r + r (1 - d r): three VALU instructions instead of the correctly
// rounded division sequence.  The PCR diagonals are O(1) (-I minus a
// diffusion operator), far from the denormal range.
__device__ __forceinline__ float rcp_nr(float d) {
    const float r = __builtin_amdgcn_rcpf(d);
    return fmaf(fmaf(-d, r, 1.0f), r, r);
}

// PCR on the 64 lane rows (row k couples to rows k -+ 1 through L and U;
// L = 0 on row 0 and U = 0 on row 63): step s eliminates the couplings to
// rows k -+ s with the neighbour rows divided by their own diagonal.
__device__ __forceinline__ float solve_f32(float L, float D, float U, float r) {
#pragma unroll
    for (int s = 1; s < 64; s <<= 1) {
        const float iD = rcp_nr(D);
        const float Ln = L * iD, Un = U * iD, rn = r * iD;
        float La, Ua, ra, Lb, Ub, rb;
        pcr::pcr_nbrs(Ln, s, La, Lb);
        pcr::pcr_nbrs(Un, s, Ua, Ub);
        pcr::pcr_nbrs(rn, s, ra, rb);
        const float nD = D - L * Ua - U * Lb;
        const float nr = r - L * ra - U * rb;
        L = -L * La;
        U = -U * Ub;
        D = nD;
        r = nr;
    }
    return r * rcp_nr(D);
}
// two right-hand sides of one matrix
__device__ __forceinline__ void solve2_f32(float L, float D, float U, float &r1, float &r2) {
#pragma unroll
    for (int s = 1; s < 64; s <<= 1) {
        const float iD = rcp_nr(D);
        const float Ln = L * iD, Un = U * iD, an = r1 * iD, bn = r2 * iD;
        float La, Ua, aa, ba, Lb, Ub, ab, bb;
        pcr::pcr_nbrs(Ln, s, La, Lb);
        pcr::pcr_nbrs(Un, s, Ua, Ub);
        pcr::pcr_nbrs(an, s, aa, ab);
        pcr::pcr_nbrs(bn, s, ba, bb);
        const float nD = D - L * Ua - U * Lb;
        r1 = r1 - L * aa - U * ab;
        r2 = r2 - L * ba - U * bb;
        L = -L * La;
        U = -U * Ub;
        D = nD;
    }
    const float iD = rcp_nr(D);
    r1 *= iD;
    r2 *= iD;
}
__device__ __forceinline__ float solve_f64(double Ld, double Dd, double Ud, double rd) {
#pragma unroll
    for (int s = 1; s < 64; s <<= 1) {
        const double iD = ca_drcp(Dd);
        const double Ln = Ld * iD, Un = Ud * iD, rn = rd * iD;
        double La, Ua, ra, Lb, Ub, rb;
        pcr::pcr_nbrs(Ln, s, La, Lb);
        pcr::pcr_nbrs(Un, s, Ua, Ub);
        pcr::pcr_nbrs(rn, s, ra, rb);
        const double nD = Dd - Ld * Ua - Ud * Lb;
        const double nr = rd - Ld * ra - Ud * rb;
        Ld = -Ld * La;
        Ud = -Ud * Ub;
        Dd = nD;
        rd = nr;
    }
    return (float)(rd * ca_drcp(Dd));
}

// Per-lane inputs of the approximate arrowhead solve (lane k = centre k and
// face k): the centre rows A(rho, u3) = (a0, a1), A(rho e, u3) = (e0, e1),
// the diffusion blocks A_ee = (elo, ed, eup) and A_hh = (hlo, hd, hup)
// (identity rows beyond the column), and the scaled residuals b.
struct Centre {
    float a0, a1, e0, e1, elo, ed, eup, hlo, hd, hup;
    float br, bh1, bh2, be;
};
struct Sol {
    float xr, xh1, xh2, xe, xw;
};

// ApproximateBlockArrowheadIterativeSolve (manual_sparse_jacobian.jl:538-578):
// a Schur complement reduction on (rho, rho e_tot) with A_11 = diag(-I, A_ee)
// inverted exactly, then `iters` stationary iterations from x = 0 on the
// velocity complement S = A_22 - A_21 inv(A_11) A_12, preconditioned by the
// complement with diag(A_ee) and solved u_h first, then u_3; finally
// x_1 = inv(A_11) (b_1 - A_12 x_2).  fr is face k's row (identity beyond
// face nl), bw0 its residual.
__device__ __forceinline__ Sol arrowhead_solve(int k, int nl, bool hasC, bool intF, bool isF, const Centre &c,
                                               const FaceRow &fr, float bw0, int iters) {
    // y_e = inv(A_ee) b_e; b_w' = b_w - A_wrho y_rho - A_we y_e (y_rho = -b_rho)
    const float ye = solve_f32(c.elo, c.ed, c.eup, c.be);
    const float pye = pcr::up1(ye), pbr = pcr::up1(c.br);
    const float pa0 = pcr::up1(c.a0), pa1 = pcr::up1(c.a1), pe0 = pcr::up1(c.e0), pe1 = pcr::up1(c.e1);
    const float ped = pcr::up1(c.ed);
    float bw = bw0, Pl = fr.lo, Pd = fr.d, Pu = fr.up;
    if (intF) {
        bw = bw0 + fr.r0 * pbr + fr.r1 * c.br - (fr.e0 * pye + fr.e1 * ye);
        const float iped = rcp_nr(ped), ied = rcp_nr(c.ed);
        Pl = fr.lo + fr.r0 * pa0 - fr.e0 * pe0 * iped;
        Pd = fr.d + fr.r0 * pa1 + fr.r1 * c.a0 - (fr.e0 * pe1 * iped + fr.e1 * c.e0 * ied);
        Pu = fr.up + fr.r1 * c.a1 - fr.e1 * c.e1 * ied;
    }
    if (!isF) {
        Pl = 0.0f;
        Pd = 1.0f;
        Pu = 0.0f;
        bw = 0.0f;
    }
    float xh1 = 0.0f, xh2 = 0.0f, xw = 0.0f;
    for (int it = 0; it < iters; ++it) {
        float rh1 = c.bh1, rh2 = c.bh2, rw = bw;
        if (it > 0) {  // r = b' - S x
            const float x1m = pcr::up1(xh1), x1p = pcr::dn1(xh1);
            const float x2m = pcr::up1(xh2), x2p = pcr::dn1(xh2);
            const float xwm = pcr::up1(xw), xwp = pcr::dn1(xw);
            if (hasC) {
                rh1 -= c.hd * xh1 + (k >= 1 ? c.hlo * x1m : 0.0f) + (k + 1 < nl ? c.hup * x1p : 0.0f);
                rh2 -= c.hd * xh2 + (k >= 1 ? c.hlo * x2m : 0.0f) + (k + 1 < nl ? c.hup * x2p : 0.0f);
            }
            const float gg = hasC ? c.e0 * xw + c.e1 * xwp : 0.0f;  // A_ew x_w
            const float mm = solve_f32(c.elo, c.ed, c.eup, gg);
            const float pmm = pcr::up1(mm);
            const float pA = c.a0 * xw + c.a1 * xwp;  // A_rho,w x_w at centre k
            const float ppA = pcr::up1(pA);
            float sx = fr.d * xw + (k >= 1 ? fr.lo * xwm : 0.0f) + (k + 1 <= nl ? fr.up * xwp : 0.0f);
            if (intF) {
                sx += fr.h0a * x1m + fr.h0b * x2m + fr.h1a * xh1 + fr.h1b * xh2;
                sx += fr.r0 * ppA + fr.r1 * pA;
                sx -= fr.e0 * pmm + fr.e1 * mm;
            }
            if (isF) rw = bw - sx;
        }
        // z_h = inv(A_hh) r_h (two components); z_w = inv(P_ww) (r_w - A_wh z_h)
        solve2_f32(c.hlo, c.hd, c.hup, rh1, rh2);
        const float z1m = pcr::up1(rh1), z2m = pcr::up1(rh2);
        if (intF) rw -= fr.h0a * z1m + fr.h0b * z2m + fr.h1a * rh1 + fr.h1b * rh2;
        const float zw = solve_f64(Pl, Pd, Pu, rw);
        if (hasC) {
            xh1 += rh1;
            xh2 += rh2;
        }
        if (isF) xw += zw;
    }
    const float xwp = pcr::dn1(xw);
    const float gg = hasC ? c.be - (c.e0 * xw + c.e1 * xwp) : 0.0f;
    Sol x;
    x.xe = solve_f32(c.elo, c.ed, c.eup, gg);
    x.xr = c.a0 * xw + c.a1 * xwp - c.br;
    x.xh1 = xh1;
    x.xh2 = xh2;
    x.xw = xw;
    return x;
}

}  // namespace idp

// V4: 16-byte row accesses in phases 1 and 3 (ncol a multiple of 4 and every
// plane 16-byte aligned; the host checks).  Compiled for 3 wavefronts per
// SIMD (at most 168 VGPRs: 167 and no spill with V4; the 4-byte fallback
// spills 36-52 bytes).
template <int NF, int NI, bool V4>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3, 8))) k_newton_idiff_pcr(DevGeom g, DevNum n, StageSrc src, float dtg, float *U, float *FI) {
    constexpr int TC = 16, R = TC + 1;
    __shared__ float tU[5][64][R];
    __shared__ float tF[5][64][R];
    const int nl = g.nlev;
    const int64_t nC = g.nC, nc = g.ncol;
    const int col0 = ca_xcd_tile(blockIdx.x, gridDim.x) * TC;
    const int tid = threadIdx.x;
    const int nf = NF >= 0 ? NF : src.nfull;
    const int ni = NI >= 0 ? NI : src.nfi;

    // ---- phase 1: U~ = sum cf full + sum ci F_I over all five components
    // (F_I carries u_h here), u_3 zero on the boundary faces
    if constexpr (V4) {
        for (int idx = tid; idx < 16 * TC; idx += 256) {
            const int r = idx / (TC / 4), c4 = 4 * (idx - r * (TC / 4));
            const int col = col0 + c4;
            const int64_t q = (int64_t)r * nc + col;
            const bool ok = col < nc;
#pragma unroll
            for (int comp = 0; comp < 5; ++comp) {
                const bool in = ok && (comp < 4 ? r < nl : (r >= 1 && r <= nl - 1));
                float4 v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                if (in) {
                    const int64_t o = comp * nC + q;
#pragma unroll
                    for (int t = 0; t < (NF >= 0 ? NF : CA_MAX_SRC); ++t)
                        if (t < nf) ca_fma4(src.cf[t], ca_ld4(src.full[t] + o), v);
#pragma unroll
                    for (int t = 0; t < (NI >= 0 ? NI : CA_MAX_SRC); ++t)
                        if (t < ni) ca_fma4(src.ci[t], ca_ld4(src.fi[t] + o), v);
                }
                tU[comp][r][c4] = v.x;
                tU[comp][r][c4 + 1] = v.y;
                tU[comp][r][c4 + 2] = v.z;
                tU[comp][r][c4 + 3] = v.w;
            }
        }
    } else
    for (int idx = tid; idx < 64 * TC; idx += 256) {
        const int r = idx / TC, cc = idx - r * TC;
        const int col = col0 + cc;
        const int64_t q = (int64_t)r * nc + col;
        const bool ok = col < nc;
#pragma unroll
        for (int comp = 0; comp < 5; ++comp) {
            const bool in = ok && (comp < 4 ? r < nl : (r >= 1 && r <= nl - 1));
            float v = 0.0f;
            if (in) {
                const int64_t o = comp * nC + q;
#pragma unroll
                for (int t = 0; t < (NF >= 0 ? NF : CA_MAX_SRC); ++t)
                    if (t < nf) v = fmaf(src.cf[t], src.full[t][o], v);
#pragma unroll
                for (int t = 0; t < (NI >= 0 ? NI : CA_MAX_SRC); ++t)
                    if (t < ni) v = fmaf(src.ci[t], src.fi[t][o], v);
            }
            tU[comp][r][cc] = v;
        }
    }
    __syncthreads();

    // ---- phase 2: one column per wavefront
    const int k = tid & 63;
    const int wid = tid >> 6;
    const bool hasC = k < nl;
    const bool intF = k >= 1 && k <= nl - 1;
    const bool isF = k <= nl;
    const int kc = hasC ? k : nl - 1;
    const int kf = isF ? k : nl;
    const float g33 = g.g33f[kf];
    const float g33p = g.g33f[kc + 1];
    const float kcv = n.cp_d / n.cv_d;
    const float idtg = 1.0f / dtg;
    for (int cc = wid; cc < TC; cc += 4) {
        const int col = col0 + cc;
        if (col >= nc) break;  // uniform across the wavefront
        const float rho = tU[0][kc][cc], u1 = tU[1][kc][cc], u2 = tU[2][kc][cc], rhoe = tU[3][kc][cc];
        const float u3 = intF ? tU[4][k][cc] : 0.0f;
        const float u3p = pcr::dn1(u3);
        const float u3m = pcr::up1(u3);
        const CV c = cv_make(g, n, kc, col, rho, u1, u2, rhoe, u3, hasC ? u3p : 0.0f);
        const float rJ = c.rho * c.J;
        const float rJm = pcr::up1(rJ), hm = pcr::up1(c.h);
        const float rJp = pcr::dn1(rJ), hp = pcr::dn1(c.h);
        // centre row k: A(rho,u3), A(rhoe,u3) and the advective residuals
        float a0 = 0.0f, a1 = 0.0f, e0 = 0.0f, e1 = 0.0f, Rr = 0.0f, Te = 0.0f;
        if (hasC) {
            float Fm = 0.0f, Fp = 0.0f, Hm = 0.0f, Hp = 0.0f;
            if (k >= 1) {
                const float rJf = 0.5f * (rJm + rJ);
                a0 = dtg * (rJf * c.iJ) * g33;
                e0 = a0 * (0.5f * (hm + c.h));
                Fm = rJf * (u3 * g33);
                Hm = Fm * (0.5f * (hm + c.h));
            }
            if (k + 1 <= nl - 1) {
                const float rJf = 0.5f * (rJ + rJp);
                a1 = dtg * (-rJf * c.iJ) * g33p;
                e1 = a1 * (0.5f * (c.h + hp));
                Fp = rJf * (u3p * g33p);
                Hp = Fp * (0.5f * (c.h + hp));
            }
            Rr = dtg * (-(Fp - Fm) * c.iJ);
            Te = -(Hp - Hm) * c.iJ;
        }
        // centre k - 1 (lane k - 1)
        CV a;
        a.rho = pcr::up1(c.rho);
        a.u1 = pcr::up1(c.u1);
        a.u2 = pcr::up1(c.u2);
        a.K = pcr::up1(c.K);
        a.Td = pcr::up1(c.Td);
        a.pd = pcr::up1(c.pd);
        a.Pid = pcr::up1(c.Pid);
        a.T = (float)a.Td;
        a.p = (float)a.pd;
        a.Pi = (float)a.Pid;
        a.ct1 = pcr::up1(c.ct1);
        a.ct2 = pcr::up1(c.ct2);
        // vertical diffusion: fluxes through face k, the tendency of centre k
        // the centre's diffusivity, and the lowest centre's wind speed (lane 0)
        const float spd0 = n.vdk == 1 ? pcr::lane0(sqrtf(c.u1 * c.ct1 + c.u2 * c.ct2)) : 0.0f;
        const float Kd = ca_Kc(g, n, col, kc, c.p, spd0);
        const float Kdm = pcr::up1(Kd), Kdp = pcr::dn1(Kd);
        float Fs = 0.0f, Fu0 = 0.0f, Fu1 = 0.0f, Fu2 = 0.0f;
        if (intF) vdiff_face(g, n, col, k, a, c, Kdm, Kd, Fs, Fu0, Fu1, Fu2);
        const float Fs1 = pcr::dn1(Fs), Fu01 = pcr::dn1(Fu0), Fu11 = pcr::dn1(Fu1), Fu21 = pcr::dn1(Fu2);
        float Rh1 = 0.0f, Rh2 = 0.0f;
        if (hasC) {
            const float iJc = c.iJ;
            Te -= (Fs1 - Fs) * iJc;
            const float d0 = (Fu01 - Fu0) * iJc, d1 = (Fu11 - Fu1) * iJc, d2 = (Fu21 - Fu2) * iJc;
            float c1, c2;
            if (n.sphere) {
                const float *B = g.basis;
                const float sk = g.sc[kc];
                c1 = sk * (d0 * B[6 * nc + col] + d1 * B[7 * nc + col] + d2 * B[8 * nc + col]);
                c2 = sk * (d0 * B[9 * nc + col] + d1 * B[10 * nc + col] + d2 * B[11 * nc + col]);
            } else {
                c1 = d0;
                c2 = d1;
            }
            const float irho = idp::rcp_nr(c.rho);
            Rh1 = dtg * (0.0f - c1 * irho);
            Rh2 = dtg * (0.0f - c2 * irho);
        }
        const float Re = dtg * Te;
        // diffusion blocks of centre k (rows >= nl: identity)
        float elo = 0.0f, ed = 1.0f, eup = 0.0f, hlo = 0.0f, hd = 1.0f, hup = 0.0f;
        if (hasC) {
            const float rm = a.rho, rp = pcr::dn1(c.rho), r0 = c.rho;
            const float iJc = idp::rcp_nr(c.J);
            const float irm = idp::rcp_nr(rm), ir0 = idp::rcp_nr(r0), irp = idp::rcp_nr(rp);
            float dlo = 0.0f, dup = 0.0f;
            if (k >= 1) {
                const float sf = g.sf[k];
                dlo = (g.Jh[col] * sf * sf * g.Jvf[k]) * g.g33f[k] * ca_rhoK(rm, r0, Kdm, Kd) * iJc;
            }
            if (k + 1 <= nl - 1) {
                const float sf = g.sf[k + 1];
                dup = (g.Jh[col] * sf * sf * g.Jvf[k + 1]) * g.g33f[k + 1] * ca_rhoK(r0, rp, Kd, Kdp) * iJc;
            }
            const float dd = -(dlo + dup);
            elo = k >= 1 ? dtg * dlo * (kcv * irm) : 0.0f;
            ed = dtg * dd * (kcv * ir0) - 1.0f;
            eup = k + 1 <= nl - 1 ? dtg * dup * (kcv * irp) : 0.0f;
            hlo = dtg * ir0 * dlo;
            hd = dtg * ir0 * dd - 1.0f;
            hup = dtg * ir0 * dup;
        }
        // face row k and its residual
        FaceRow fr;
        fr.r0 = fr.r1 = fr.e0 = fr.e1 = fr.h0a = fr.h0b = fr.h1a = fr.h1b = fr.lo = fr.up = 0.0f;
        fr.d = 1.0f;
        float Ru3 = 0.0f;
        if (intF) {
            fr = jac_face_row(g, n, k, col, a, c, u3m, u3, u3p, dtg);
            Ru3 = dtg * (pgf_u3(g, n, k, a, c) - beta_w(g, n, k, col) * u3);
        } else if (isF) {
            fr.d = -dtg * beta_w(g, n, k, col) - 1.0f;
        }
        // ---- ApproximateBlockArrowheadIterativeSolve
        idp::Centre cb;
        cb.a0 = a0;
        cb.a1 = a1;
        cb.e0 = e0;
        cb.e1 = e1;
        cb.elo = hasC ? elo : 0.0f;
        cb.ed = hasC ? ed : 1.0f;
        cb.eup = hasC ? eup : 0.0f;
        cb.hlo = hasC ? hlo : 0.0f;
        cb.hd = hasC ? hd : 1.0f;
        cb.hup = hasC ? hup : 0.0f;
        cb.br = Rr;
        cb.be = hasC ? Re : 0.0f;
        cb.bh1 = hasC ? Rh1 : 0.0f;
        cb.bh2 = hasC ? Rh2 : 0.0f;
        const idp::Sol x = idp::arrowhead_solve(k, nl, hasC, intF, isF, cb, fr, Ru3, n.siters);
        const float xh1 = x.xh1, xh2 = x.xh2, xe = x.xe, xw = x.xw, xr = x.xr;
        // Newton update
        const float rho_n = rho - xr;
        const float u1_n = u1 - xh1, u2_n = u2 - xh2;
        float rhoe_n = rhoe - xe;
        const float u3_n = intF ? u3 - xw : 0.0f;
        // post-Newton upwind correction at the new state
        if (n.upw != 0) {
            const float u3_np = pcr::dn1(u3_n);
            const float hn = cv_enthalpy(g, n, kc, col, rho_n, u1_n, u2_n, rhoe_n, u3_n, hasC ? u3_np : 0.0f);
            const float href = pcr::lane0(hn);
            const float hr = hn - href;
            const float sk = g.sc[kc];
            const float rJn = rho_n * (g.Jh[col] * sk * sk * g.Jvc[kc]);
            const float h2 = pcr::up(hr, 2), h1 = pcr::up1(hr), h4 = pcr::dn1(hr);
            const float rJn1 = pcr::up1(rJn);
            float dF = 0.0f;
            if (intF) {
                const float v = u3_n * g33;
                const float rJf = 0.5f * (rJn1 + rJn);
                const float upw = upwind_product(n.upw, k, nl, v, n.dt, h2, h1, hr, h4);
                const float cen = upwind_product(0, k, nl, v, n.dt, 0.0f, h1, hr, 0.0f);
                dF = rJf * upw - rJf * cen;
            }
            const float dF1 = pcr::dn1(dF);
            const float corr = hasC ? -(dF1 - dF) * (g.iJh[col] * g.isJc[kc]) : 0.0f;
            rhoe_n += dtg * corr;
        }
        // U and F_I = (U - U~) / dtg
        if (hasC) {
            tU[0][k][cc] = rho_n;
            tU[1][k][cc] = u1_n;
            tU[2][k][cc] = u2_n;
            tU[3][k][cc] = rhoe_n;
            tF[0][k][cc] = (rho_n - rho) * idtg;
            tF[1][k][cc] = (u1_n - u1) * idtg;
            tF[2][k][cc] = (u2_n - u2) * idtg;
            tF[3][k][cc] = (rhoe_n - rhoe) * idtg;
        }
        if (isF) {
            tU[4][k][cc] = u3_n;
            tF[4][k][cc] = (u3_n - u3) * idtg;
        }
    }
    __syncthreads();

    // ---- phase 3: coalesced stores of U and F_I
    if constexpr (V4) {
        for (int idx = tid; idx < 16 * TC; idx += 256) {
            const int r = idx / (TC / 4), c4 = 4 * (idx - r * (TC / 4));
            const int col = col0 + c4;
            if (col >= nc) continue;
            const int64_t q = (int64_t)r * nc + col;
            auto row4 = [&](const float (*t)[R]) {
                return make_float4(t[r][c4], t[r][c4 + 1], t[r][c4 + 2], t[r][c4 + 3]);
            };
            if (r < nl) {
#pragma unroll
                for (int comp = 0; comp < 4; ++comp) {
                    *reinterpret_cast<float4 *>(U + comp * nC + q) = row4(tU[comp]);
                    if (FI) *reinterpret_cast<float4 *>(FI + comp * nC + q) = row4(tF[comp]);
                }
            }
            if (r <= nl) {
                *reinterpret_cast<float4 *>(U + 4 * nC + q) = row4(tU[4]);
                if (FI) *reinterpret_cast<float4 *>(FI + 4 * nC + q) = row4(tF[4]);
            }
        }
        return;
    }
    for (int idx = tid; idx < 64 * TC; idx += 256) {
        const int r = idx / TC, cc = idx - r * TC;
        const int col = col0 + cc;
        if (col >= nc) continue;
        const int64_t q = (int64_t)r * nc + col;
        if (r < nl) {
#pragma unroll
            for (int comp = 0; comp < 4; ++comp) {
                U[comp * nC + q] = tU[comp][r][cc];
                if (FI) FI[comp * nC + q] = tF[comp][r][cc];
            }
        }
        if (r <= nl) {
            U[4 * nC + q] = tU[4][r][cc];
            if (FI) FI[4 * nC + q] = tF[4][r][cc];
        }
    }
}

#define CA_INST_IDIFF(NF, NI, V4)                                                                          \
    template __global__ void k_newton_idiff_pcr<NF, NI, V4>(DevGeom, DevNum, StageSrc, float, float *, float *);
#define CA_INST_IDIFF_ALL(V4) \
    CA_INST_IDIFF(2, 0, V4)   \
    CA_INST_IDIFF(3, 1, V4)   \
    CA_INST_IDIFF(4, 2, V4)   \
    CA_INST_IDIFF(-1, -1, V4)
CA_INST_IDIFF_ALL(false)
CA_INST_IDIFF_ALL(true)

// ldiv! with implicit diffusion from the stored bands (the hook entry point
// ca_ldiv after ca_update_jacobian): dY = solve(A, rscale R) with the
// arrowhead solve of the fused stage, one wavefront per column.  A workgroup
// stages 8 columns of the 26 band and residual planes through LDS with
// coalesced row loads; the results leave through the residual tile.
__global__ void __launch_bounds__(256) k_ldiv_approx_pcr(DevGeom g, Bands A, DiffBands E, int iters, float rscale,
                                                         float *dY, const float *R) {
    constexpr int TC = 8, RW = TC + 1;
    // planes: 0-3 centre rows (Aru0, Aru1, Aeu0, Aeu1); 4-14 face rows (Aur0, Aur1,
    // Aue0, Aue1, Auh0a, Auh0b, Auh1a, Auh1b, lo, d, up); 15-20 E (elo, ed, eup,
    // hlo, hd, hup); 21-25 R (rho, u_1, u_2, rho e_tot, u_3), then dY in place
    constexpr int NP = 26;
    __shared__ float t[NP][64][RW];
    const int nl = g.nlev;
    const int64_t nC = g.nC, nc = g.ncol;
    const int col0 = ca_xcd_tile(blockIdx.x, gridDim.x) * TC;
    const int tid = threadIdx.x;
    const float *src[NP] = {A.Aru0, A.Aru1, A.Aeu0, A.Aeu1, A.Aur0, A.Aur1, A.Aue0, A.Aue1, A.Auh0a,
                            A.Auh0b, A.Auh1a, A.Auh1b, A.lo, A.d, A.up, E.elo, E.ed, E.eup, E.hlo,
                            E.hd, E.hup, R, R + nC, R + 2 * nC, R + 3 * nC, R + 4 * nC};
    for (int idx = tid; idx < 64 * TC; idx += 256) {
        const int r = idx / TC, cc = idx - r * TC;
        const int col = col0 + cc;
        const int64_t q = (int64_t)r * nc + col;
        const bool ok = col < nc, cen = r < nl, face = r <= nl;
#pragma unroll
        for (int p = 0; p < NP; ++p) {
            const bool fp = (p >= 4 && p <= 14) || p == 25;
            t[p][r][cc] = ok && (fp ? face : cen) ? src[p][q] : 0.0f;
        }
    }
    __syncthreads();
    const int k = tid & 63, wid = tid >> 6;
    const bool hasC = k < nl, intF = k >= 1 && k <= nl - 1, isF = k <= nl;
    for (int cc = wid; cc < TC; cc += 4) {
        if (col0 + cc >= nc) break;  // uniform across the wavefront
        idp::Centre c;
        c.a0 = t[0][k][cc];
        c.a1 = t[1][k][cc];
        c.e0 = t[2][k][cc];
        c.e1 = t[3][k][cc];
        c.elo = hasC ? t[15][k][cc] : 0.0f;
        c.ed = hasC ? t[16][k][cc] : 1.0f;
        c.eup = hasC ? t[17][k][cc] : 0.0f;
        c.hlo = hasC ? t[18][k][cc] : 0.0f;
        c.hd = hasC ? t[19][k][cc] : 1.0f;
        c.hup = hasC ? t[20][k][cc] : 0.0f;
        c.br = rscale * t[21][k][cc];
        c.bh1 = rscale * t[22][k][cc];
        c.bh2 = rscale * t[23][k][cc];
        c.be = rscale * t[24][k][cc];
        FaceRow fr;
        fr.r0 = t[4][k][cc];
        fr.r1 = t[5][k][cc];
        fr.e0 = t[6][k][cc];
        fr.e1 = t[7][k][cc];
        fr.h0a = t[8][k][cc];
        fr.h0b = t[9][k][cc];
        fr.h1a = t[10][k][cc];
        fr.h1b = t[11][k][cc];
        fr.lo = t[12][k][cc];
        fr.d = isF ? t[13][k][cc] : 1.0f;
        fr.up = t[14][k][cc];
        const float bw = rscale * t[25][k][cc];
        const idp::Sol x = idp::arrowhead_solve(k, nl, hasC, intF, isF, c, fr, bw, iters);
        if (hasC) {
            t[21][k][cc] = x.xr;
            t[22][k][cc] = x.xh1;
            t[23][k][cc] = x.xh2;
            t[24][k][cc] = x.xe;
        }
        if (isF) t[25][k][cc] = x.xw;
    }
    __syncthreads();
    for (int idx = tid; idx < 64 * TC; idx += 256) {
        const int r = idx / TC, cc = idx - r * TC;
        const int col = col0 + cc;
        if (col >= nc) continue;
        const int64_t q = (int64_t)r * nc + col;
        if (r < nl)
#pragma unroll
            for (int p = 0; p < 4; ++p) dY[p * nC + q] = t[21 + p][r][cc];
        if (r <= nl) dY[4 * nC + q] = t[25][r][cc];
    }
}

// U~ = sum cf full + sum ci fi over all five components (the implicit
// tendencies carry u_h here), u_3 zero on the boundary faces
__global__ void __launch_bounds__(256) k_stage_sum(DevGeom g, StageSrc src, float *out) {
    const int64_t S = 4 * g.nC + g.nF;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= S) return;
    if (i >= 4 * g.nC) {
        const int64_t f = (i - 4 * g.nC) / g.ncol;
        if (f == 0 || f == g.nlev) {
            out[i] = 0.0f;
            return;
        }
    }
    float s = 0.0f;
    for (int t = 0; t < src.nfull; ++t) s = fmaf(src.cf[t], src.full[t][i], s);
    for (int t = 0; t < src.nfi; ++t) s = fmaf(src.ci[t], src.fi[t][i], s);
    out[i] = s;
}

// U = U~ - dY (the Newton update), u_3 zero on the boundary faces
__global__ void __launch_bounds__(256) k_newton_update(DevGeom g, const float *Ut, const float *dY, float *U) {
    const int64_t S = 4 * g.nC + g.nF;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= S) return;
    float v = Ut[i] - dY[i];
    if (i >= 4 * g.nC) {
        const int64_t f = (i - 4 * g.nC) / g.ncol;
        if (f == 0 || f == g.nlev) v = 0.0f;
    }
    U[i] = v;
}

// the post-Newton correction U.rho e_tot += dtg C.rho e_tot (C may be null),
// then F_I = (U - U~) / dtg over all five components
__global__ void __launch_bounds__(256) k_stage_finish(DevGeom g, const float *Ut, const float *C, float dtg, float *U,
                                                      float *FI) {
    const int64_t S = 4 * g.nC + g.nF;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= S) return;
    float u = U[i];
    if (C && i >= 3 * g.nC && i < 4 * g.nC) {
        u += dtg * C[i];
        U[i] = u;
    }
    FI[i] = (u - Ut[i]) / dtg;
}
