// ca_kernels.h -- kernel declarations and launch-argument structures.
#pragma once
#include "ca_device.h"

// centre quantities of one column level (recomputed in registers, never stored)
struct CV {
    float rho, u1, u2, rhoe;
    float ct1, ct2;  // contravariant horizontal velocity
    float K, T, h, p, Pi, J, iJ;  // iJ = 1 / J
    double Td, pd, Pid;  // temperature, pressure and Exner function in float64
};

// Jacobian bands (update_jacobian! storage for the hook-level ldiv)
struct Bands {
    float *Aru0, *Aru1, *Aeu0, *Aeu1;                 // centre rows
    float *Aur0, *Aur1, *Aue0, *Aue1;                 // face rows: centres f-1, f
    float *Auh0a, *Auh0b, *Auh1a, *Auh1b;
    float *lo, *d, *up;
};

// implicit-diffusion blocks (tridiagonal on centres): (rho e_tot, rho e_tot)
// and (u_h, u_h)
struct DiffBands {
    float *elo, *ed, *eup, *hlo, *hd, *hup;
};

#define CA_MAX_SRC 8
// stage-state assembly: U~ = sum cf[t] * full[t] + sum ci[t] * fi[t]
// (fi sources hold the rho, rho e_tot and u_3 planes of an implicit tendency)
struct StageSrc {
    const float *full[CA_MAX_SRC];
    float cf[CA_MAX_SRC];
    const float *fi[CA_MAX_SRC];
    float ci[CA_MAX_SRC];
    int nfull, nfi;
};

struct FinalSrc {
    const float *fe[CA_MAX_SRC];
    float be[CA_MAX_SRC];
    const float *fi[CA_MAX_SRC];
    float bi[CA_MAX_SRC];
    int nfe, nfi;
    float dt;
    int fi_full;  // 1: the fi sources hold all five components (implicit diffusion)
};

// column kernels (ca_column.hip)
__global__ void k_precompute(DevGeom g, DevNum n, float *Y, float *K, float *T, float *p, float *h);
__global__ void k_implicit_tendency(DevGeom g, DevNum n, float *Yt, const float *Y);
__global__ void k_update_jacobian(DevGeom g, DevNum n, const float *Y, float dtg, Bands B);
__global__ void k_ldiv(DevGeom g, Bands B, float *dY, const float *R, float *scr);
__global__ void k_correct(DevGeom g, DevNum n, float *Yt, const float *Y);
__global__ void k_newton_stage(DevGeom g, DevNum n, StageSrc src, float dtg, float *U, float *FI, float *scr);
__global__ void k_ldiv_approx_pcr(DevGeom g, Bands A, DiffBands E, int iters, float rscale, float *dY,
                                  const float *R);
template <int NF, int NI, bool V4>
__global__ void k_newton_idiff_pcr(DevGeom g, DevNum n, StageSrc src, float dtg, float *U, float *FI);
template <int TC, int NF, int NI, bool V4>
__global__ void k_newton_pcr(DevGeom g, DevNum n, StageSrc src, float dtg, float *U, float *FI, float *Ut);

// The final combination of an ARS343 step fused into its last explicit
// tendency.  The implicit tableau is stiffly accurate (b_j = a_I[4][j]) and
// F_I,j = (U_j - U~_j) / (dt a_I[j][j]), so
//   u + dt sum_j b_j (F_E,j + F_I,j) = U_4 + dt sum_j (b_j - a_E[4][j]) F_E,j
// with a_E[4][4] = 0: the last stage's tendency need not be stored, nor the
// implicit ones read.
struct FinFuse {
    const float *fe[3];  // F_E of stages 1..3
    float c[3];          // dt (b_j - a_E[4][j])
    float cg;            // dt b_4, the weight of the tendency being computed
};

// explicit tendency (ca_spectral.hip)
__global__ void k_laplacians(DevGeom g, DevNum n, const float *Y, float *L, int chunk);
// FIN: Yt receives U + the combination of FinFuse instead of the tendency at U
template <bool SPH, bool HYPER, bool VDIFF, bool FIN>
__global__ void k_explicit(DevGeom g, DevNum n, const float *Y, const float *L, float *Yt, int chunk,
                           FinFuse fin);

// DSS (ca_dss.hip): out-of-place weighted average over shared nodes.
// NC centre planes, the vector pair at planes (VEC, VEC + 1) (-1: none), a
// face plane after them if FACE.
template <int NC, int VEC, bool FACE>
__global__ void k_dss_gather(DevGeom g, const float *src, float *dst, int filter, int chunk);
// the same, staged through LDS per element cluster
template <int NC, int VEC, bool FACE, bool V4>
__global__ void k_dss_lds(DevGeom g, const float *src, float *dst, int filter, int chunk);
__global__ void k_combine(DevGeom g, const float *u, float *out, FinalSrc s);
