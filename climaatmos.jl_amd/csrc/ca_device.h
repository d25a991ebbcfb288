// ca_device.h -- device-side data structures and pointwise physics of the
// dry-dycore step (gfx950, float32).
//
// Layout: every field is level-major, value(k, col) at [k * ncol + col];
// col = 16 * element + 4 * j + i.  Geometry is factored: per-column
// horizontal metric planes and per-level vertical tables.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ca_dss_table.h"

#define CA_NQ 4
#define CA_NN 16

// GLL Nq = 4: quadrature weight products W[4 j + i] = w_i w_j and the
// derivative matrix D[i][k] = l_k'(xi_i).  Filled once by ca_create.
__constant__ float c_W[CA_NN];
__constant__ float c_iW[CA_NN];  // 1 / W
__constant__ float c_D[CA_NQ][CA_NQ];

struct DevGeom {
    int nelem, ncol, nlev;
    int64_t nC, nF;  // centre / face plane sizes
    // per-column planes (ncol each)
    const float *Jh, *gu11, *gu12, *gu22, *gd11, *gd12, *gd22, *f3n, *f12a, *f12b;
    const float *iJh;                    // 1 / J_h
    const float *basis;  // 12 planes: a^1 (x,y,z), a^2, a_1, a_2
    // per-level tables
    const float *zc, *Jvc, *sc, *dzc;    // centres [nlev]; dzc = Jvc
    const float *zf, *Jvf, *sf, *dPhif;  // faces [nlev + 1]; dPhif[f] = g (zc[f] - zc[f-1])
    const float *zeta_c, *zeta_f;        // Rayleigh profiles sin^2(...) (0 below zd)
    const float *Phic;                   // g zc
    // reciprocals (per level, computed in double): 1 / s_c^2, 1 / Jv_c,
    // 1 / (s_c^2 Jv_c); g^33 = 1 / Jv_f^2, 1 / Jv_f, 1 / s_f, 1 / (s_f^2 Jv_f)
    const float *is2c, *iJvc, *isJc;
    const float *g33f, *iJvf, *isf, *isJf;
    const double *dPhif_d;               // dPhif in float64
    const double *Phic_d;                // g zc in float64
    // per-column gather table of the weighted DSS, one row per column, the
    // copies of the column's node in the group's canonical order, padded to
    // four slots with (own column, weight 0):
    //   gtc[col]  int4    column of each slot
    //   gtw[col]  float4  weight WJ / sum WJ of each slot
    //   gtM[col]  float4 x 4, slot t: w_t (a_1 . a^1, a_1 . a^2, a_2 . a^1, a_2 . a^2)
    //             with a_i of the own column and a^j of the slot's column
    // An interior node is its own single slot of weight 1.
    const int4 *gtc;
    const float4 *gtw, *gtM;
    // LDS-staged form of the same table.  Elements are grouped into clusters
    // of at most CA_DSS_OWN elements; a cluster's element list holds its own
    // elements followed by every other element one of their nodes' copies
    // lives in (at most CA_DSS_EMAX in all).  Per (cluster, thread): the own
    // column (-1: none), the four slots as indices into the cluster's
    // staged values (list position * 16 + node), the weights and transforms.
    int ncl;
    const int *dl_list, *dl_nE, *dl_col;
    const ushort4 *dl_slot;
    const float4 *dl_w, *dl_M;
    const float *aw_col, *D0_col;        // optional per-column sponge alpha_w / diffusivity
    const float *kdec;                   // [nlev] exp(-(z_c - z_sfc) / H) of DecayWithHeightDiffusion
    const float *Tsfc;                   // per-column surface temperature (surface fluxes)
};

struct DevNum {
    float R_d, cp_d, cv_d, T_0, p0inv, kex, kap, grav, Tsr, Tmr, Tmin;
    double dR_d, dcp_d, dcv_d, dT_0, dp0inv, dkex, dTsr, dTmr, dTmin;
    double dicv_d;  // 1 / cv_d
    double dkex_r;  // kex - 2/7 (|.| <= 1e-6, checked at create)
    float nu_v, nu_s, dd;
    float alpha_uh, alpha_w, dt;
    float Cd, Ch, gust;  // surface flux exchange coefficients and wind-speed floor
    int hyper, upw, rayleigh, vdiff, sphere, sflux;
    int idiff, siters;  // vertical diffusion implicit; iterations of its linear solve
    int vdk;            // closure: 0 DecayWithHeightDiffusion, 1 VerticalDiffusion
    float CE, za;       // VerticalDiffusion's C_E and half the lowest layer's depth
};

// ------------------------------------------------------------- eddy diffusivity
// K_h of the centre at level k of column col (ᶜcompute_eddy_diffusivity_coefficient,
// cache/eddy_diffusivity_coefficient.jl:16-41; precomputed_quantities.jl:651-676):
// DecayWithHeightDiffusion D0 exp(-(z - z_sfc) / H), or VerticalDiffusion
// C_E |u_h| z_a from the lowest centre's wind speed spd0, tapered by
// exp(-((85000 - p) / 10000)^2) above 850 hPa.
__device__ __forceinline__ float ca_Kc(const DevGeom &g, const DevNum &n, int col, int k, float p, float spd0) {
    if (n.vdk == 0) return g.D0_col[col] * g.kdec[k];
    const float KE = n.CE * spd0 * n.za;
    if (p > 85000.0f) return KE;
    const float x = (85000.0f - p) * 1.0e-4f;
    return KE * expf(-(x * x));
}
// rho K on a face: interp(rho) / interp(1 / max(K, eps)), the harmonic mean
// of the two centre values (vertical_diffusion_boundary_layer.jl:86-90)
__device__ __forceinline__ float ca_rhoK(float ra, float rb, float Ka, float Kb) {
    const float e = 1.1920929e-07f;
    const float Kf = 1.0f / (0.5f * (1.0f / fmaxf(Ka, e) + 1.0f / fmaxf(Kb, e)));
    return (0.5f * (ra + rb)) * Kf;
}

// ------------------------------------------------------------- launch order
// Workgroups are dispatched to the 8 XCDs round-robin (workgroup b on XCD
// b mod 8), each XCD with its own L2.  ca_xcd_tile maps workgroup b of nb to
// a tile index so that every XCD takes one contiguous range of tiles, in
// order: tiles that share cache lines (narrow column tiles of a level-major
// plane) are then fetched through the same L2.  A bijection of [0, nb).
__device__ __forceinline__ int ca_xcd_tile(int b, int nb) {
    const int q = nb >> 3, r = nb & 7, x = b & 7, j = b >> 3;
    return x < r ? x * (q + 1) + j : r * (q + 1) + (x - r) * q + j;
}

// ------------------------------------------------------------- thermodynamics
// Dry air (Thermodynamics 1.3.0 as used at precomputed_quantities.jl:786-815):
// e_int = cv_d (T - T_0) - R_d T_0, p = rho R_d T, h_tot = e_tot + R_d T,
// Pi = (p / p_ref)^(R_d / cp_d); T is floored at T_min_sgs (:786-796).
__device__ __forceinline__ float ca_temperature(const DevNum &n, float e_int) {
    float T = n.T_0 + (e_int + n.R_d * n.T_0) / n.cv_d;
    return fmaxf(T, n.Tmin);
}
// Exner function (p / p0)^kex through the hardware log2 / exp2 (p > 0)
__device__ __forceinline__ float ca_exner(const DevNum &n, float p) {
    return __builtin_amdgcn_exp2f(n.kex * __builtin_amdgcn_logf(p * n.p0inv));
}
// 1 / x to about 1 ulp (v_rcp_f32)
__device__ __forceinline__ float ca_rcp(float x) { return __builtin_amdgcn_rcpf(x); }
// 1 / x in float64 (finite, normal x): v_rcp_f64 and two Newton-Raphson steps
__device__ __forceinline__ double ca_drcp(double x) {
    double r = __builtin_amdgcn_rcp(x);
    r = fma(fma(-x, r, 1.0), r, r);
    r = fma(fma(-x, r, 1.0), r, r);
    return r;
}
// Exner function q^kex in float64 for q = p / p0 > 0, without a float64
// log or exp.  The dry-air exponent is 2/7 up to kex_r = kex - 2/7 (rounding
// of the parameter set's kappa_d; |kex_r| <= 1e-6 is checked at create):
//   y0 = q^(2/7) from the hardware log2 / exp2 (relative error ~2e-7);
//   one Newton step on y^7 = q^2: y = (6 y0 + q^2 / y0^6) / 7 (relative
//   error ~3 e0^2 ~ 1e-13);
//   q^kex = y q^kex_r = y (1 + kex_r ln q + O((kex_r ln q)^2)), where
//   (kex_r ln q)^2 <= 1e-10 for q >= 1e-3 and 0 for an exact 2/7.
__device__ __forceinline__ double ca_dexner(const DevNum &n, double q) {
    const float qf = (float)q;
    const float l2 = __builtin_amdgcn_logf(qf);  // log2 q
    const double y0 = (double)__builtin_amdgcn_exp2f((2.0f / 7.0f) * l2);
    const double y2 = y0 * y0;
    const double y6 = y2 * y2 * y2;
    const double y = (6.0 * y0 + (q * q) * ca_drcp(y6)) * (1.0 / 7.0);
    return fma(y, n.dkex_r * (0.6931471805599453 * (double)l2), y);
}
// reference temperature T_r(Pi) = T_min + (T_sfc - T_min) Pi^7 (refstate_thermodynamics.jl:80-88)
__device__ __forceinline__ float ca_pow7(float x) {
    float x2 = x * x;
    float x4 = x2 * x2;
    return x4 * x2 * x;
}
__device__ __forceinline__ float ca_Tref(const DevNum &n, float Pi) { return n.Tmr + (n.Tsr - n.Tmr) * ca_pow7(Pi); }
// Phi_r(p_b) - Phi_r(p_a) in a cancellation-free form:
//   -cp_d [T_min ln(Pi_b / Pi_a) + (T_sfc - T_min) / 7 (Pi_b^7 - Pi_a^7)]
// with ln(Pi_b / Pi_a) = kex ln(p_b / p_a) and Pi_b^7 - Pi_a^7 = Pi_a^7 expm1(7 ln(Pi_b / Pi_a)).
// ipa = 1 / pa
__device__ __forceinline__ float ca_dPhiref(const DevNum &n, float pa, float ipa, float Pia, float pb) {
    float lr = n.kex * log1pf((pb - pa) * ipa);
    return -n.cp_d * (n.Tmr * lr + (n.Tsr - n.Tmr) * (1.0f / 7.0f) * ca_pow7(Pia) * expm1f(7.0f * lr));
}
// Both differences from one log1p and one expm1: with x = expm1(lr),
// Pi_b - Pi_a = Pi_a x and expm1(7 lr) = (1 + x)^7 - 1 expanded in x (no
// cancellation for small x).
__device__ __forceinline__ void ca_dref(const DevNum &n, float pa, float ipa, float Pia, float pb, float &dPhr,
                                        float &dPi) {
    const float lr = n.kex * log1pf((pb - pa) * ipa);
    const float x = expm1f(lr);
    const float e7 = x * (7.0f + x * (21.0f + x * (35.0f + x * (35.0f + x * (21.0f + x * (7.0f + x))))));
    dPi = Pia * x;
    dPhr = -n.cp_d * (n.Tmr * lr + (n.Tsr - n.Tmr) * (1.0f / 7.0f) * ca_pow7(Pia) * e7);
}

// horizontal metric at scale s (no topography): contravariant u^a = g^{ab} u_b / s^2
__device__ __forceinline__ void ca_contra(const DevGeom &g, int col, float is2, float a1, float a2, float &o1,
                                          float &o2) {
    float g11 = g.gu11[col], g12 = g.gu12[col], g22 = g.gu22[col];
    o1 = is2 * (g11 * a1 + g12 * a2);
    o2 = is2 * (g12 * a1 + g22 * a2);
}
__device__ __forceinline__ void ca_cov(const DevGeom &g, int col, float s2, float a1, float a2, float &o1,
                                       float &o2) {
    float g11 = g.gd11[col], g12 = g.gd12[col], g22 = g.gd22[col];
    o1 = s2 * (g11 * a1 + g12 * a2);
    o2 = s2 * (g12 * a1 + g22 * a2);
}

// Access `off` bytes past a wavefront-uniform base pointer.  With the lane's
// part a 32-bit byte offset the address is a scalar base plus a vector
// offset: no per-lane 64-bit address arithmetic per access.
__device__ __forceinline__ float ca_ldo(const float *base, uint32_t off) {
    return *reinterpret_cast<const float *>(reinterpret_cast<const char *>(base) + off);
}
__device__ __forceinline__ void ca_sto(float *base, uint32_t off, float v) {
    *reinterpret_cast<float *>(reinterpret_cast<char *>(base) + off) = v;
}

// The horizontal metric of one column (level independent), loaded once by
// kernels that loop over levels, so the per-level arithmetic reads registers.
struct ColMet {
    float gu11, gu12, gu22, gd11, gd12, gd22;
};
__device__ __forceinline__ ColMet ca_colmet(const DevGeom &g, int col) {
    ColMet c;
    c.gu11 = g.gu11[col];
    c.gu12 = g.gu12[col];
    c.gu22 = g.gu22[col];
    c.gd11 = g.gd11[col];
    c.gd12 = g.gd12[col];
    c.gd22 = g.gd22[col];
    return c;
}
__device__ __forceinline__ void ca_contra(const ColMet &c, float is2, float a1, float a2, float &o1, float &o2) {
    o1 = is2 * (c.gu11 * a1 + c.gu12 * a2);
    o2 = is2 * (c.gu12 * a1 + c.gu22 * a2);
}
__device__ __forceinline__ void ca_cov(const ColMet &c, float s2, float a1, float a2, float &o1, float &o2) {
    o1 = s2 * (c.gd11 * a1 + c.gd12 * a2);
    o2 = s2 * (c.gd12 * a1 + c.gd22 * a2);
}

// ------------------------------------------------- element-local derivatives
// The 16 nodes of an element occupy one 16-lane DPP row of a wavefront
// (lane = 16 e + 4 j + i).  Along the first reference coordinate the four
// values of a node's line are quad broadcasts (DPP quad_perm); along the
// second they are row rotations by 4, 8, 12 lanes (DPP row_ror), each node
// weighting a rotation with its own coefficient of D.  Both are plain VALU
// moves: no LDS traffic, no permute latency.
struct Nod {
    int base, i, j;  // lane of node (0, 0) of the element; node indices
    float c1[4];     // D[i][m]            (strong, first coordinate)
    float c1t[4];    // D[m][i]            (weak / transpose)
    float c2[4];     // D[j][(j - k) & 3]  (strong, second coordinate, rotation 4k)
    float c2t[4];    // D[(j - k) & 3][j]
};
__device__ __forceinline__ Nod ca_node(int lane) {
    Nod q;
    q.base = lane & ~15;
    q.i = lane & 3;
    q.j = (lane >> 2) & 3;
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        q.c1[m] = c_D[q.i][m];
        q.c1t[m] = c_D[m][q.i];
        q.c2[m] = c_D[q.j][(q.j - m) & 3];
        q.c2t[m] = c_D[(q.j - m) & 3][q.j];
    }
    return q;
}
// Every pattern used here (quad_perm, row_ror, row_newbcast with full row and
// bank masks) has a valid source lane for every lane, so bound_ctrl changes
// no value; it lets the compiler drop the zero "old" operand it would
// otherwise materialise before each move.
template <int CTRL>
__device__ __forceinline__ float ca_dpp(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, true));
}
// value at node (0, 0) of the element (DPP row_newbcast:0)
__device__ __forceinline__ float ca_elem0(float v) { return ca_dpp<0x150>(v); }
// quad broadcast of element m: quad_perm(m, m, m, m)
#define CA_QB(m) ((m) | ((m) << 2) | ((m) << 4) | ((m) << 6))
// row rotate right by r lanes (within the 16-lane row)
#define CA_ROR(r) (0x120 + (r))
// Each derivative is one block of four VALU instructions: a multiply and three
// v_fmac_f32_dpp, each moving f by its DPP pattern in the operand fetch.  The
// compiler folds a DPP move into v_mul / v_add but not into v_fmac (its
// accumulator is tied to the destination), which would cost a separate
// v_mov_b32_dpp per term, and these operators are VALU-issue bound.  Every
// product and sum rounds as in fmaf(c[m], f(m), r) chained from m = 0 to 3.
// The s_nop 1 ahead of the block is the two wait states a DPP read of a VGPR
// needs after a VALU write of it (the compiler's hazard recognizer does not
// look inside inline assembly); f is not written inside the block, so one
// suffices.  The result is early-clobber: it is written before f and the
// coefficients are last read.
#define CA_DPP_MASKS " row_mask:0xf bank_mask:0xf bound_ctrl:1"
#define CA_QUAD(m) " quad_perm:[" #m "," #m "," #m "," #m "]" CA_DPP_MASKS "\n\t"
// sum_m D[i][m] f(m, j)   (c = q.c1) and sum_m D[m][i] g(m, j)   (c = q.c1t)
__device__ __forceinline__ float ca_dquad(const float c[4], float f) {
    float r;
    asm("s_nop 1\n\t"
        "v_mul_f32_dpp %0, %1, %2" CA_QUAD(0)
        "v_fmac_f32_dpp %0, %1, %3" CA_QUAD(1)
        "v_fmac_f32_dpp %0, %1, %4" CA_QUAD(2)
        "v_fmac_f32_dpp %0, %1, %5" CA_QUAD(3)
        : "=&v"(r)
        : "v"(f), "v"(c[0]), "v"(c[1]), "v"(c[2]), "v"(c[3]));
    return r;
}
// sum_k c[k] f(i, (j - k) & 3): rotation of the row by 4k lanes
__device__ __forceinline__ float ca_dror(const float c[4], float f) {
    float r;
    asm("s_nop 1\n\t"
        "v_mul_f32_e32 %0, %1, %2\n\t"
        "v_fmac_f32_dpp %0, %1, %3 row_ror:4" CA_DPP_MASKS "\n\t"
        "v_fmac_f32_dpp %0, %1, %4 row_ror:8" CA_DPP_MASKS "\n\t"
        "v_fmac_f32_dpp %0, %1, %5 row_ror:12" CA_DPP_MASKS
        : "=&v"(r)
        : "v"(f), "v"(c[0]), "v"(c[1]), "v"(c[2]), "v"(c[3]));
    return r;
}
#undef CA_QUAD
#undef CA_DPP_MASKS
// sum_m D[i][m] f(m, j)
__device__ __forceinline__ float ca_d1(const Nod &q, float f) { return ca_dquad(q.c1, f); }
// sum_m D[m][i] g(m, j)   (transpose: weak forms)
__device__ __forceinline__ float ca_d1t(const Nod &q, float f) { return ca_dquad(q.c1t, f); }
// sum_m D[j][m] f(i, m): rotation by 4k brings f(i, (j - k) & 3)
__device__ __forceinline__ float ca_d2(const Nod &q, float f) { return ca_dror(q.c2, f); }
__device__ __forceinline__ float ca_d2t(const Nod &q, float f) { return ca_dror(q.c2t, f); }

// Metric factors of one node at one level: W = quadrature weight product,
// J = 3-D Jacobian, and the reciprocals the operators multiply by.
struct Met {
    float W, iW, J, iJ, WJ, iWJ;
};
__device__ __forceinline__ Met ca_met(float W, float iW, float J, float iJ) {
    Met m;
    m.W = W;
    m.iW = iW;
    m.J = J;
    m.iJ = iJ;
    m.WJ = W * J;
    m.iWJ = iW * iJ;
    return m;
}

// Spectral-element operators at one level (docs/src/discretization.md:72-135).
// strong gradient -> C12
__device__ __forceinline__ void op_grad(const Nod &q, float f, float &o1, float &o2) {
    o1 = ca_d1(q, f);
    o2 = ca_d2(q, f);
}
// strong divergence of contravariant (v1, v2): (D (J v1) + D (J v2)) / J
__device__ __forceinline__ float op_div(const Nod &q, const Met &m, float v1, float v2) {
    return (ca_d1(q, m.J * v1) + ca_d2(q, m.J * v2)) * m.iJ;
}
// weak divergence: -(D^T (W J v1) + D^T (W J v2)) / (W J)
__device__ __forceinline__ float op_wdiv(const Nod &q, const Met &m, float v1, float v2) {
    return -(ca_d1t(q, m.WJ * v1) + ca_d2t(q, m.WJ * v2)) * m.iWJ;
}
// weak gradient -> C12: -(D^T (W f)) / W
__device__ __forceinline__ void op_wgrad(const Nod &q, const Met &m, float f, float &o1, float &o2) {
    const float Wf = m.W * f;
    o1 = -ca_d1t(q, Wf) * m.iW;
    o2 = -ca_d2t(q, Wf) * m.iW;
}
// strong curl of C12 (u1, u2) -> CT3
__device__ __forceinline__ float op_curl3(const Nod &q, const Met &m, float u1, float u2) {
    return (ca_d1(q, u2) - ca_d2(q, u1)) * m.iJ;
}
// weak curl of C12 -> CT3: (-D^T (W u2) + D^T (W u1)) / (W J)
__device__ __forceinline__ float op_wcurl3(const Nod &q, const Met &m, float u1, float u2) {
    return (-ca_d1t(q, m.W * u2) + ca_d2t(q, m.W * u1)) * m.iWJ;
}
// strong curl of C3 u3 -> CT12
__device__ __forceinline__ void op_curl12(const Nod &q, const Met &m, float u3, float &o1, float &o2) {
    o1 = ca_d2(q, u3) * m.iJ;
    o2 = -ca_d1(q, u3) * m.iJ;
}
// weak curl of C3 -> CT12
__device__ __forceinline__ void op_wcurl12(const Nod &q, const Met &m, float u3, float &o1, float &o2) {
    const float Wu = m.W * u3;
    o1 = -ca_d2t(q, Wu) * m.iWJ;
    o2 = ca_d1t(q, Wu) * m.iWJ;
}

// ----------------------------------------------------------- Laplacians
// prep_hyperdiffusion_tendency! (hyperdiffusion.jl:141-147, dry) at one point.
// The four Laplacian planes at one (node, level): lanes 16 r .. 16 r + 15
// hold the 16 nodes of one element at one level (DPP rows); cm, Jh, iJh:
// the column's horizontal metric; u3a, u3b: u_3 at the faces below and above.
// The s_d - s_d,r field is taken relative to its value at the element's first
// node (a constant inside the element, annihilated by the gradient), which
// removes the large geopotential part.
__device__ __forceinline__ void lap_point(const DevGeom &g, const DevNum &n, const Nod &q, float W, float iW,
                                          float Jh, float iJh, const ColMet &cm, int k, float rho, float u1,
                                          float u2, float rhoe, float u3a, float u3b, float out[4]) {
    const float s = g.sc[k], s2 = s * s, is2 = g.is2c[k], Jv = g.Jvc[k];
    const Met m = ca_met(W, iW, Jh * s2 * Jv, iJh * g.isJc[k]);
    float ct1, ct2;
    ca_contra(cm, is2, u1, u2, ct1, ct2);
    const float ga = g.g33f[k], gb = g.g33f[k + 1];
    const float K = 0.5f * ((u1 * ct1 + u2 * ct2) + 0.5f * (u3a * ga * u3a + u3b * gb * u3b));
    const double e_tot_d = (double)rhoe * ca_drcp((double)rho);
    const double Td = fmax(n.dT_0 + (e_tot_d - (double)K - g.Phic_d[k] + n.dR_d * n.dT_0) * n.dicv_d, n.dTmin);
    const float T = (float)Td;
    const float p = (float)((double)rho * n.dR_d * Td);
    const float Pi = ca_exner(n, p);
    const float cu3 = 0.5f * (u3a + u3b);
    // vector Laplacian of c_u
    const float dv = op_div(q, m, ct1, ct2);
    float gd1, gd2;
    op_wgrad(q, m, dv, gd1, gd2);
    const float cr3 = op_curl3(q, m, u1, u2);
    float cr1, cr2;
    op_curl12(q, m, cu3, cr1, cr2);
    float cc1, cc2;
    ca_cov(cm, s2, cr1, cr2, cc1, cc2);
    const float cc3 = Jv * Jv * cr3;
    const float w3 = op_wcurl3(q, m, cc1, cc2);
    float w1, w2;
    op_wcurl12(q, m, cc3, w1, w2);
    float wc1, wc2;
    ca_cov(cm, s2, w1, w2, wc1, wc2);
    const float wc3 = Jv * Jv * w3;
    out[0] = gd1 - wc1;
    out[1] = gd2 - wc2;
    out[2] = 0.0f - wc3;
    // scalar Laplacian of the dry static energy deviation
    const float p0 = ca_elem0(p), Pi0 = ca_elem0(Pi);
    const float psi = n.cp_d * (T - ca_Tref(n, Pi)) - ca_dPhiref(n, p0, ca_rcp(p0), Pi0, p);
    float q1, q2;
    op_grad(q, psi, q1, q2);
    float v1, v2;
    ca_contra(cm, is2, q1, q2, v1, v2);
    out[3] = op_wdiv(q, m, v1, v2);
}

// van Leer (Lin et al. 1994) slope with the monotone local-extrema
// constraint, on the upwind cell a1 with neighbours a0, a2.
__device__ __forceinline__ float ca_vl_slope(float a0, float a1, float a2) {
    float d = 0.5f * ((a1 - a0) + (a2 - a1));
    float mn = fminf(fminf(a0, a1), a2), mx = fmaxf(fmaxf(a0, a1), a2);
    float lim = fminf(fabsf(d), fminf(2.0f * (a1 - mn), 2.0f * (mx - a1)));
    return d > 0.0f ? lim : (d < 0.0f ? -lim : 0.0f);
}

#define CA_CHECK_LAUNCH() (void)0
