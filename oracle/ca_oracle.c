/*
 * ca_oracle.c -- CPU restatement of one ClimaAtmos dry-dycore IMEX step.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the native HIP library and, when
 * built with -DORACLE_F32 -fopenmp, the CPU timing baseline.  The product
 * never links it.
 *
 * The restatement follows the reference operator by operator, materialising
 * every field the reference materialises, in the same order:
 *   - explicit tendency    src/prognostic_equations/remaining_tendency.jl:48-58
 *   - horizontal dynamics  src/prognostic_equations/advection.jl:36-91
 *   - hyperdiffusion       src/prognostic_equations/hyperdiffusion.jl:116-147, 247-307
 *   - vertical advection   src/prognostic_equations/advection.jl:205-290
 *   - implicit tendency    src/prognostic_equations/implicit/implicit_tendency.jl:185-298
 *   - Jacobian + solve     src/prognostic_equations/implicit/manual_sparse_jacobian.jl:713-870,579-584
 *   - post-Newton upwind   src/prognostic_equations/implicit/implicit_tendency.jl:322-339
 *   - precomputed fields   src/cache/precomputed_quantities.jl:698-831
 *   - reference state      src/utils/refstate_thermodynamics.jl:58-166
 *   - operators            src/utils/abbreviations.jl:72-256 (ClimaCore semantics)
 *   - DSS                  src/prognostic_equations/constrain_state.jl:59-64
 *   - IMEX driver          src/simulation/integrator.jl:193-225 (ARS343, 1 Newton iteration)
 * Spectral-element operators use the GLL Nq=4 collocation of
 * docs/src/discretization.md:72-135; the split divergence is the formula of
 * docs/src/discretization.md:248-253.
 */
#include "ca_oracle.h"

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif
#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef ora_real real;

#define NQ 4
#define NN 16

struct ora_ctx {
    ora_geom g;
    ora_num n;
    int nl, ncol;
    int64_t nC, nF, S;     /* centre/face plane sizes, state size */
    real *zc, *zf, *Jvc, *Jvf, *sc, *sf, *Phic;
    real *Jc, *Jf;         /* 3-D Jacobians (k, col) */
    real W[NN], D[NQ][NQ];
    /* precomputed */
    real *cu3, *K, *T, *p, *h, *fu3;
    real *uh0;             /* horizontal wind speed of the lowest centre, per column */
    /* hyperdiffusion Laplacians */
    real *L1, *L2, *L3, *Ls;
    /* Jacobian bands (per column, level-major) */
    real *Aru0, *Aru1, *Aeu0, *Aeu1;              /* centre k: faces k, k+1 */
    real *Aur0, *Aur1, *Aue0, *Aue1;              /* face f: centres f-1, f */
    real *Auh0a, *Auh0b, *Auh1a, *Auh1b;          /* face f: (c=f-1)(1,2), (c=f)(1,2) */
    real *Auu_lo, *Auu_d, *Auu_up;
    /* implicit diffusion: (rho e_tot, rho e_tot) and (u_h, u_h), tridiagonal on centres */
    real *Aee_lo, *Aee_d, *Aee_up, *Ahh_lo, *Ahh_d, *Ahh_up;
    /* scratch */
    real *t[12];
    real *FE[4], *FI[4], *Ut, *U, *Rv, *dY;
    double dtgamma;
    int round_storage;     /* ora_set_storage_rounding */
};

#define IDX(k, col) ((int64_t)(k) * c->ncol + (col))

/* The Float32 build is the CPU timing baseline and runs its column, element
 * and node-group loops on OpenMP threads (every iteration writes only its
 * own column / element / group); the Float64 checker is single-threaded. */
#ifdef _OPENMP
#define ORA_PAR_FOR _Pragma("omp parallel for schedule(static)")
#else
#define ORA_PAR_FOR
#endif

static void *zalloc(size_t n) { return calloc(n ? n : 1, 1); }

int64_t ora_state_size(const ora_ctx *c) { return c->S; }

/* ---------------------------------------------------------------- setup */

int ora_create(const ora_geom *g, const ora_num *n, ora_ctx **out) {
    ora_ctx *c = (ora_ctx *)zalloc(sizeof(ora_ctx));
    if (!c) return -1;
    c->g = *g;
    c->n = *n;
    c->nl = g->nlev;
    c->ncol = g->ncol;
    c->nC = (int64_t)g->nlev * g->ncol;
    c->nF = (int64_t)(g->nlev + 1) * g->ncol;
    c->S = 4 * c->nC + c->nF;
    int nl = c->nl;
    c->zc = zalloc(sizeof(real) * nl);
    c->zf = zalloc(sizeof(real) * (nl + 1));
    c->Jvc = zalloc(sizeof(real) * nl);
    c->Jvf = zalloc(sizeof(real) * (nl + 1));
    c->sc = zalloc(sizeof(real) * nl);
    c->sf = zalloc(sizeof(real) * (nl + 1));
    c->Phic = zalloc(sizeof(real) * nl);
    for (int f = 0; f <= nl; ++f) c->zf[f] = (real)g->z_face[f];
    for (int k = 0; k < nl; ++k) {
        c->zc[k] = (real)(0.5 * (g->z_face[k] + g->z_face[k + 1]));
        c->Jvc[k] = (real)(g->z_face[k + 1] - g->z_face[k]);
    }
    c->Jvf[0] = (real)(g->z_face[1] - g->z_face[0]);
    c->Jvf[nl] = (real)(g->z_face[nl] - g->z_face[nl - 1]);
    for (int f = 1; f < nl; ++f) c->Jvf[f] = c->zc[f] - c->zc[f - 1];
    for (int k = 0; k < nl; ++k)
        c->sc[k] = g->radius > 0 ? (real)((g->radius + c->zc[k]) / g->radius) : 1;
    for (int f = 0; f <= nl; ++f)
        c->sf[f] = g->radius > 0 ? (real)((g->radius + c->zf[f]) / g->radius) : 1;
    for (int k = 0; k < nl; ++k) c->Phic[k] = (real)n->grav * c->zc[k];
    c->Jc = zalloc(sizeof(real) * c->nC);
    c->Jf = zalloc(sizeof(real) * c->nF);
    for (int k = 0; k < nl; ++k)
        for (int i = 0; i < c->ncol; ++i)
            c->Jc[IDX(k, i)] = (real)g->J_h[i] * c->sc[k] * c->sc[k] * c->Jvc[k];
    for (int f = 0; f <= nl; ++f)
        for (int i = 0; i < c->ncol; ++i)
            c->Jf[IDX(f, i)] = (real)g->J_h[i] * c->sf[f] * c->sf[f] * c->Jvf[f];
    for (int j = 0; j < NQ; ++j)
        for (int i = 0; i < NQ; ++i) c->W[4 * j + i] = (real)(g->gll_w[i] * g->gll_w[j]);
    for (int i = 0; i < NQ; ++i)
        for (int k = 0; k < NQ; ++k) c->D[i][k] = (real)g->gll_D[4 * i + k];
    real **cen[] = {&c->cu3, &c->K, &c->T, &c->p, &c->h, &c->L1, &c->L2, &c->L3, &c->Ls,
                    &c->Aru0, &c->Aru1, &c->Aeu0, &c->Aeu1, &c->Aee_lo, &c->Aee_d, &c->Aee_up,
                    &c->Ahh_lo, &c->Ahh_d, &c->Ahh_up};
    for (size_t q = 0; q < sizeof(cen) / sizeof(cen[0]); ++q) *cen[q] = zalloc(sizeof(real) * c->nC);
    real **fac[] = {&c->fu3, &c->Aur0, &c->Aur1, &c->Aue0, &c->Aue1, &c->Auh0a, &c->Auh0b,
                    &c->Auh1a, &c->Auh1b, &c->Auu_lo, &c->Auu_d, &c->Auu_up};
    for (size_t q = 0; q < sizeof(fac) / sizeof(fac[0]); ++q) *fac[q] = zalloc(sizeof(real) * c->nF);
    for (int q = 0; q < 12; ++q) c->t[q] = zalloc(sizeof(real) * c->nF);
    for (int q = 0; q < 4; ++q) {
        c->FE[q] = zalloc(sizeof(real) * c->S);
        c->FI[q] = zalloc(sizeof(real) * c->S);
    }
    c->Ut = zalloc(sizeof(real) * c->S);
    c->U = zalloc(sizeof(real) * c->S);
    c->Rv = zalloc(sizeof(real) * c->S);
    c->dY = zalloc(sizeof(real) * c->S);
    c->uh0 = zalloc(sizeof(real) * c->ncol);
    *out = c;
    return 0;
}

void ora_destroy(ora_ctx *c) {
    if (!c) return;
    real *ptrs[] = {c->zc, c->zf, c->Jvc, c->Jvf, c->sc, c->sf, c->Phic, c->Jc, c->Jf,
                    c->cu3, c->K, c->T, c->p, c->h, c->fu3, c->L1, c->L2, c->L3, c->Ls,
                    c->Aru0, c->Aru1, c->Aeu0, c->Aeu1, c->Aur0, c->Aur1, c->Aue0, c->Aue1,
                    c->Auh0a, c->Auh0b, c->Auh1a, c->Auh1b, c->Auu_lo, c->Auu_d, c->Auu_up,
                    c->Aee_lo, c->Aee_d, c->Aee_up, c->Ahh_lo, c->Ahh_d, c->Ahh_up,
                    c->Ut, c->U, c->Rv, c->dY, c->uh0};
    for (size_t q = 0; q < sizeof(ptrs) / sizeof(ptrs[0]); ++q) free(ptrs[q]);
    for (int q = 0; q < 12; ++q) free(c->t[q]);
    for (int q = 0; q < 4; ++q) {
        free(c->FE[q]);
        free(c->FI[q]);
    }
    free(c);
}

/* state component pointers */
#define RHO(Y) (Y)
#define UH1(Y) ((Y) + c->nC)
#define UH2(Y) ((Y) + 2 * c->nC)
#define RHOE(Y) ((Y) + 3 * c->nC)
#define U3(Y) ((Y) + 4 * c->nC)

/* -------------------------------------------------- thermodynamics (dry) */
/* Dry-air relations of Thermodynamics 1.3.0 as used at
 * precomputed_quantities.jl:786-815 and pinned by the Jacobian algebra of
 * manual_sparse_jacobian.jl:816-825: e_int = cv_d (T - T_0) - R_d T_0,
 * p = rho R_d T, h_tot = e_tot + R_d T, Pi = (p / p_ref_theta)^(R_d / cp_d). */
static inline real exner(const ora_ctx *c, real p) {
    return (real)pow((double)p / c->n.p_ref_theta, c->n.R_d / c->n.cp_d);
}
/* refstate_thermodynamics.jl:80-88 (T_r), :99-103 (theta_vr), :123-131 (Phi_r),
 * :161-166 (sd_r); s_ref = 7 (:22) */
static inline real T_ref(const ora_ctx *c, real Pi) {
    real Pi7 = Pi * Pi * Pi * Pi * Pi * Pi * Pi;
    return (real)c->n.T_min_ref + (real)(c->n.T_surf_ref - c->n.T_min_ref) * Pi7;
}
static inline real Phi_ref(const ora_ctx *c, real Pi) {
    real Pi7 = Pi * Pi * Pi * Pi * Pi * Pi * Pi;
    return -(real)c->n.cp_d * ((real)c->n.T_min_ref * (real)log((double)Pi) +
                               (real)(c->n.T_surf_ref - c->n.T_min_ref) / 7 * (Pi7 - 1));
}

/* ------------------------------------------------ horizontal SE operators */
/* One level, all elements.  Node n = 4 j + i inside an element, i along the
 * first reference coordinate.  J is the 3-D Jacobian at the level. */
static void op_grad(const ora_ctx *c, const real *f, real *o1, real *o2) {
    ORA_PAR_FOR
    for (int e = 0; e < c->g.nelem; ++e) {
        const real *fe = f + NN * e;
        for (int j = 0; j < NQ; ++j)
            for (int i = 0; i < NQ; ++i) {
                real a = 0, b = 0;
                for (int m = 0; m < NQ; ++m) {
                    a += c->D[i][m] * fe[4 * j + m];
                    b += c->D[j][m] * fe[4 * m + i];
                }
                o1[NN * e + 4 * j + i] = a;
                o2[NN * e + 4 * j + i] = b;
            }
    }
}
static void op_div(const ora_ctx *c, const real *J, const real *v1, const real *v2, real *o) {
    ORA_PAR_FOR
    for (int e = 0; e < c->g.nelem; ++e) {
        real a1[NN], a2[NN];
        for (int q = 0; q < NN; ++q) {
            a1[q] = J[NN * e + q] * v1[NN * e + q];
            a2[q] = J[NN * e + q] * v2[NN * e + q];
        }
        for (int j = 0; j < NQ; ++j)
            for (int i = 0; i < NQ; ++i) {
                real s = 0;
                for (int m = 0; m < NQ; ++m) s += c->D[i][m] * a1[4 * j + m] + c->D[j][m] * a2[4 * m + i];
                o[NN * e + 4 * j + i] = s / J[NN * e + 4 * j + i];
            }
    }
}
static void op_wdiv(const ora_ctx *c, const real *J, const real *v1, const real *v2, real *o) {
    ORA_PAR_FOR
    for (int e = 0; e < c->g.nelem; ++e) {
        real a1[NN], a2[NN];
        for (int q = 0; q < NN; ++q) {
            a1[q] = c->W[q] * J[NN * e + q] * v1[NN * e + q];
            a2[q] = c->W[q] * J[NN * e + q] * v2[NN * e + q];
        }
        for (int j = 0; j < NQ; ++j)
            for (int i = 0; i < NQ; ++i) {
                real s = 0;
                for (int m = 0; m < NQ; ++m) s += c->D[m][i] * a1[4 * j + m] + c->D[m][j] * a2[4 * m + i];
                int q = 4 * j + i;
                o[NN * e + q] = -s / (c->W[q] * J[NN * e + q]);
            }
    }
}
static void op_wgrad(const ora_ctx *c, const real *f, real *o1, real *o2) {
    ORA_PAR_FOR
    for (int e = 0; e < c->g.nelem; ++e) {
        real a[NN];
        for (int q = 0; q < NN; ++q) a[q] = c->W[q] * f[NN * e + q];
        for (int j = 0; j < NQ; ++j)
            for (int i = 0; i < NQ; ++i) {
                real s1 = 0, s2 = 0;
                for (int m = 0; m < NQ; ++m) {
                    s1 += c->D[m][i] * a[4 * j + m];
                    s2 += c->D[m][j] * a[4 * m + i];
                }
                int q = 4 * j + i;
                o1[NN * e + q] = -s1 / c->W[q];
                o2[NN * e + q] = -s2 / c->W[q];
            }
    }
}
static void op_curl3(const ora_ctx *c, const real *J, const real *u1, const real *u2, real *o) {
    ORA_PAR_FOR
    for (int e = 0; e < c->g.nelem; ++e)
        for (int j = 0; j < NQ; ++j)
            for (int i = 0; i < NQ; ++i) {
                real s = 0;
                for (int m = 0; m < NQ; ++m)
                    s += c->D[i][m] * u2[NN * e + 4 * j + m] - c->D[j][m] * u1[NN * e + 4 * m + i];
                o[NN * e + 4 * j + i] = s / J[NN * e + 4 * j + i];
            }
}
static void op_wcurl3(const ora_ctx *c, const real *J, const real *u1, const real *u2, real *o) {
    ORA_PAR_FOR
    for (int e = 0; e < c->g.nelem; ++e) {
        real a1[NN], a2[NN];
        for (int q = 0; q < NN; ++q) {
            a1[q] = c->W[q] * u1[NN * e + q];
            a2[q] = c->W[q] * u2[NN * e + q];
        }
        for (int j = 0; j < NQ; ++j)
            for (int i = 0; i < NQ; ++i) {
                real s = 0;
                for (int m = 0; m < NQ; ++m) s += -c->D[m][i] * a2[4 * j + m] + c->D[m][j] * a1[4 * m + i];
                int q = 4 * j + i;
                o[NN * e + q] = s / (c->W[q] * J[NN * e + q]);
            }
    }
}
static void op_curl12(const ora_ctx *c, const real *J, const real *u3, real *o1, real *o2) {
    ORA_PAR_FOR
    for (int e = 0; e < c->g.nelem; ++e)
        for (int j = 0; j < NQ; ++j)
            for (int i = 0; i < NQ; ++i) {
                real a = 0, b = 0;
                for (int m = 0; m < NQ; ++m) {
                    a += c->D[j][m] * u3[NN * e + 4 * m + i];
                    b += c->D[i][m] * u3[NN * e + 4 * j + m];
                }
                int q = NN * e + 4 * j + i;
                o1[q] = a / J[q];
                o2[q] = -b / J[q];
            }
}
static void op_wcurl12(const ora_ctx *c, const real *J, const real *u3, real *o1, real *o2) {
    ORA_PAR_FOR
    for (int e = 0; e < c->g.nelem; ++e) {
        real a[NN];
        for (int q = 0; q < NN; ++q) a[q] = c->W[q] * u3[NN * e + q];
        for (int j = 0; j < NQ; ++j)
            for (int i = 0; i < NQ; ++i) {
                real s1 = 0, s2 = 0;
                for (int m = 0; m < NQ; ++m) {
                    s1 += c->D[m][j] * a[4 * m + i];
                    s2 += c->D[m][i] * a[4 * j + m];
                }
                int q = 4 * j + i;
                o1[NN * e + q] = -s1 / (c->W[q] * J[NN * e + q]);
                o2[NN * e + q] = s2 / (c->W[q] * J[NN * e + q]);
            }
    }
}

/* metric conversions at scale s (no topography): contravariant horizontal
 * u^a = g^{ab}_h / s^2 u_b; covariant u_a = s^2 g_{ab,h} u^b */
static inline void to_contra(const ora_ctx *c, int col, real s, real a1, real a2, real *o1, real *o2) {
    const double *G = c->g.gup;
    real inv = 1 / (s * s);
    *o1 = inv * ((real)G[col] * a1 + (real)G[c->ncol + col] * a2);
    *o2 = inv * ((real)G[c->ncol + col] * a1 + (real)G[2 * c->ncol + col] * a2);
}
static inline void to_cov(const ora_ctx *c, int col, real s, real a1, real a2, real *o1, real *o2) {
    const double *G = c->g.gdn;
    real s2 = s * s;
    *o1 = s2 * ((real)G[col] * a1 + (real)G[c->ncol + col] * a2);
    *o2 = s2 * ((real)G[c->ncol + col] * a1 + (real)G[2 * c->ncol + col] * a2);
}

/* ------------------------------------------------------------------ DSS */
/* Weighted direct stiffness summation (constrain_state.jl:59-64): each shared
 * node's value becomes the J-weighted average over its copies; covariant
 * horizontal vectors are averaged as Cartesian vectors. */
static void dss_scalar(const ora_ctx *c, real *f, int nplanes) {
    for (int k = 0; k < nplanes; ++k) {
        real *fk = f + (int64_t)k * c->ncol;
        ORA_PAR_FOR
        for (int n = 0; n < c->g.n_dss; ++n) {
            real s = 0;
            for (int q = c->g.dss_ptr[n]; q < c->g.dss_ptr[n + 1]; ++q)
                s += (real)c->g.dss_w[q] * fk[c->g.dss_col[q]];
            for (int q = c->g.dss_ptr[n]; q < c->g.dss_ptr[n + 1]; ++q) fk[c->g.dss_col[q]] = s;
        }
    }
}
static void dss_vector(const ora_ctx *c, real *u1, real *u2, int nplanes) {
    const double *B = c->g.basis;
    int64_t nc = c->ncol;
    for (int k = 0; k < nplanes; ++k) {
        real *a = u1 + (int64_t)k * nc, *b = u2 + (int64_t)k * nc;
        ORA_PAR_FOR
        for (int n = 0; n < c->g.n_dss; ++n) {
            real X[3] = {0, 0, 0};
            for (int q = c->g.dss_ptr[n]; q < c->g.dss_ptr[n + 1]; ++q) {
                int col = c->g.dss_col[q];
                real w = (real)c->g.dss_w[q];
                for (int d = 0; d < 3; ++d)
                    X[d] += w * (a[col] * (real)B[(0 + d) * nc + col] + b[col] * (real)B[(3 + d) * nc + col]);
            }
            for (int q = c->g.dss_ptr[n]; q < c->g.dss_ptr[n + 1]; ++q) {
                int col = c->g.dss_col[q];
                real v1 = 0, v2 = 0;
                for (int d = 0; d < 3; ++d) {
                    v1 += X[d] * (real)B[(6 + d) * nc + col];
                    v2 += X[d] * (real)B[(9 + d) * nc + col];
                }
                a[col] = v1;
                b[col] = v2;
            }
        }
    }
}
static void dss_state(ora_ctx *c, real *Y) {
    if (c->g.nelem == 0) return;
    dss_scalar(c, RHO(Y), c->nl);
    dss_vector(c, UH1(Y), UH2(Y), c->nl);
    dss_scalar(c, RHOE(Y), c->nl);
    dss_scalar(c, U3(Y), c->nl + 1);
}
void ora_dss(ora_ctx *c, ora_real *Y) { dss_state(c, (real *)Y); }

/* --------------------------------------------------- precomputed fields */
/* set_implicit_precomputed_quantities!, precomputed_quantities.jl:698-831,
 * dry branch.  Without topography u_h^3 = 0, so the impenetrability filter
 * (:486-554) sets u_3 = 0 on the bottom and top faces. */
static void set_precomputed(ora_ctx *c, real *Y) {
    int nl = c->nl, nc = c->ncol;
    real *u3 = U3(Y);
    ORA_PAR_FOR
    for (int col = 0; col < nc; ++col) {
        u3[IDX(0, col)] = 0;
        u3[IDX(nl, col)] = 0;
    }
    const double *G = c->g.gup;
    for (int f = 0; f <= nl; ++f)
        ORA_PAR_FOR
        for (int col = 0; col < nc; ++col) {
            real g33 = 1 / (c->Jvf[f] * c->Jvf[f]);
            c->fu3[IDX(f, col)] = g33 * u3[IDX(f, col)];
        }
    for (int k = 0; k < nl; ++k)
        ORA_PAR_FOR
        for (int col = 0; col < nc; ++col) {
            int64_t q = IDX(k, col);
            real rho = RHO(Y)[q], u1 = UH1(Y)[q], u2 = UH2(Y)[q], rhoe = RHOE(Y)[q];
            real ua = u3[IDX(k, col)], ub = u3[IDX(k + 1, col)];
            c->cu3[q] = (ua + ub) / 2;
            real s = c->sc[k], is2 = 1 / (s * s);
            /* compute_kinetic, utilities.jl:194-212 */
            real uh_dot = is2 * ((real)G[col] * u1 * u1 + 2 * (real)G[nc + col] * u1 * u2 +
                                 (real)G[2 * nc + col] * u2 * u2);
            real ga = 1 / (c->Jvf[k] * c->Jvf[k]), gb = 1 / (c->Jvf[k + 1] * c->Jvf[k + 1]);
            real vv = (ua * ga * ua + ub * gb * ub) / 2;
            real Kk = (uh_dot + vv) / 2;
            c->K[q] = Kk;
            if (k == 0) c->uh0[col] = (real)sqrt((double)uh_dot);
            real e_tot = rhoe / rho;
            real e_int = e_tot - Kk - c->Phic[k];
            real T = (real)c->n.T_0 + (e_int + (real)(c->n.R_d * c->n.T_0)) / (real)c->n.cv_d;
            if (T < (real)c->n.T_min_sgs) T = (real)c->n.T_min_sgs;
            c->T[q] = T;
            c->h[q] = e_tot + (real)c->n.R_d * T;
            c->p[q] = rho * (real)c->n.R_d * T;
        }
}
void ora_set_precomputed(ora_ctx *c, ora_real *Y) { set_precomputed(c, (real *)Y); }

void ora_get_precomputed(const ora_ctx *c, double *K, double *T, double *p, double *h) {
    ORA_PAR_FOR
    for (int64_t q = 0; q < c->nC; ++q) {
        K[q] = c->K[q];
        T[q] = c->T[q];
        p[q] = c->p[q];
        h[q] = c->h[q];
    }
}
void ora_get_laplacians(const ora_ctx *c, double *out) {
    ORA_PAR_FOR
    for (int64_t q = 0; q < c->nC; ++q) {
        out[q] = c->L1[q];
        out[c->nC + q] = c->L2[q];
        out[2 * c->nC + q] = c->L3[q];
        out[3 * c->nC + q] = c->Ls[q];
    }
}

/* ------------------------------------------------ vertical operator pieces */
static inline real rayleigh_zeta(const ora_ctx *c, real z) {
    real zmax = c->zf[c->nl], zd = (real)c->n.zd_rayleigh;
    if (!(z > zd)) return 0;
    real s = (real)sin(M_PI * (double)((z - zd) / (zmax - zd)) / 2);
    return s * s;
}
static inline real alpha_w_of(const ora_ctx *c, int col) {
    return c->n.alpha_w_col ? (real)c->n.alpha_w_col[col] : (real)c->n.alpha_w;
}

/* face value of the upwinded product U(u3, chi) (abbreviations.jl:227-256).
 * 0: central I^f(chi) u3; 1: first-order upwind; 3: Lin (1994) van Leer with
 * the monotone local-extrema constraint and first-order one-sided stencils
 * on the two faces next to each boundary. */
static real upwind_face(const ora_ctx *c, int scheme, const real *chi, int f, int col, real v) {
    int nl = c->nl;
    real am = chi[IDX(f - 1, col)], ap = chi[IDX(f, col)];
    if (scheme == 0) return v * (am + ap) / 2;
    if (scheme == 1 || f == 1 || f == nl - 1) return v >= 0 ? v * am : v * ap;
    real dt = (real)c->n.dt;
    real cour = v * dt;
    if (v >= 0) {
        real a0 = chi[IDX(f - 2, col)], a1 = am, a2 = ap;
        real d = ((a1 - a0) + (a2 - a1)) / 2;
        real mn = fmin(fmin(a0, a1), a2), mx = fmax(fmax(a0, a1), a2);
        real lim = fmin(fabs(d), fmin(2 * (a1 - mn), 2 * (mx - a1)));
        real slope = d > 0 ? lim : (d < 0 ? -lim : 0);
        return v * (a1 + (1 - cour) * slope / 2);
    } else {
        real a0 = am, a1 = ap, a2 = chi[IDX(f + 1, col)];
        real d = ((a1 - a0) + (a2 - a1)) / 2;
        real mn = fmin(fmin(a0, a1), a2), mx = fmax(fmax(a0, a1), a2);
        real lim = fmin(fabs(d), fmin(2 * (a1 - mn), 2 * (mx - a1)));
        real slope = d > 0 ? lim : (d < 0 ? -lim : 0);
        return v * (a1 - (1 + cour) * slope / 2);
    }
}

/* vertical_transport(rho, u3, chi, dt, scheme), implicit_tendency.jl:120-143:
 * -advdiv(I^f(rho J)/J_f * U(u3, chi)) with zero flux on boundary faces. */
static void vertical_transport(const ora_ctx *c, const real *rho, const real *chi, int scheme, real *out) {
    int nl = c->nl, nc = c->ncol;
    real *F = c->t[11];
    ORA_PAR_FOR
    for (int col = 0; col < nc; ++col) {
        F[IDX(0, col)] = 0;
        F[IDX(nl, col)] = 0;
        for (int f = 1; f < nl; ++f) {
            real rJ = (rho[IDX(f - 1, col)] * c->Jc[IDX(f - 1, col)] + rho[IDX(f, col)] * c->Jc[IDX(f, col)]) / 2;
            real Jf = c->Jf[IDX(f, col)];
            real u = c->fu3[IDX(f, col)];
            F[IDX(f, col)] = Jf * (rJ / Jf * upwind_face(c, scheme, chi, f, col, u));
        }
        for (int k = 0; k < nl; ++k)
            out[IDX(k, col)] = -((F[IDX(k + 1, col)] - F[IDX(k, col)]) / c->Jc[IDX(k, col)]);
    }
}

static void vertical_diffusion_tendency(ora_ctx *c, real *Yt, real *Y);

/* ------------------------------------------------------ implicit tendency */
/* implicit_tendency!, implicit_tendency.jl:36-98 and :185-298 (dry subset) */
static void implicit_tendency(ora_ctx *c, real *Yt, real *Y) {
    int nl = c->nl, nc = c->ncol;
    memset(Yt, 0, sizeof(real) * c->S);
    real *rho = RHO(Y);
    /* rho: -advdiv(I^f(rho J)/J_f u^3)  (:202) */
    real *one = c->t[10];
    ORA_PAR_FOR
    for (int64_t q = 0; q < c->nC; ++q) one[q] = 1;
    vertical_transport(c, rho, one, 0, c->t[9]);
    /* the central transport of chi = 1 is -advdiv(I^f(rho J)/J_f u^3 * 1) */
    ORA_PAR_FOR
    for (int64_t q = 0; q < c->nC; ++q) RHO(Yt)[q] += c->t[9][q];
    /* rho e_tot: central vertical transport of h_tot (:209-210) */
    vertical_transport(c, rho, c->h, 0, c->t[9]);
    ORA_PAR_FOR
    for (int64_t q = 0; q < c->nC; ++q) RHOE(Yt)[q] += c->t[9][q];
    /* u_3: -(grad Phi - grad Phi_r + cp_d I^f(theta_v - theta_vr) grad Pi) (:287-293) */
    real *Pi = c->t[0], *Phr = c->t[1], *thd = c->t[2];
    ORA_PAR_FOR
    for (int64_t q = 0; q < c->nC; ++q) {
        real P = exner(c, c->p[q]);
        Pi[q] = P;
        Phr[q] = Phi_ref(c, P);
        thd[q] = c->T[q] / P - T_ref(c, P) / P;
    }
    real cp = (real)c->n.cp_d;
    ORA_PAR_FOR
    for (int col = 0; col < nc; ++col)
        for (int f = 1; f < nl; ++f) {
            int64_t a = IDX(f - 1, col), b = IDX(f, col);
            real gPhi = c->Phic[f] - c->Phic[f - 1];
            real gPhr = Phr[b] - Phr[a];
            real th = (thd[a] + thd[b]) / 2;
            real gPi = Pi[b] - Pi[a];
            U3(Yt)[IDX(f, col)] -= gPhi - gPhr + cp * th * gPi;
        }
    /* Rayleigh sponge on u_3 (:295-296, rayleigh_sponge.jl:46-57) */
    if (c->n.rayleigh)
        for (int f = 0; f <= nl; ++f)
            ORA_PAR_FOR
            for (int col = 0; col < nc; ++col) {
                real beta = alpha_w_of(c, col) * rayleigh_zeta(c, c->zf[f]);
                U3(Yt)[IDX(f, col)] += -beta * U3(Y)[IDX(f, col)];
            }
    /* vertical diffusion with diff_mode Implicit (:73-82) */
    if (c->n.vert_diff && c->n.implicit_diffusion) vertical_diffusion_tendency(c, Yt, Y);
}
void ora_implicit_tendency(ora_ctx *c, ora_real *Yt, ora_real *Y) { implicit_tendency(c, (real *)Yt, (real *)Y); }

/* update_diffusion_jacobian!, manual_sparse_jacobian.jl:1031-1261, dry,
 * DecayWithHeightDiffusion (K_u = K_h = D0, harmonic face mean as in the
 * tendency), for one column:
 *   D = advdiv_matrix . diag(I^f(rho) / I^f(1 / max(K, eps))) . grad_matrix   (:1067-1068)
 *   (rho e_tot, rho e_tot) = dtg D diag(cp_d / (cv_d rho)) - I                   (:1106-1108, :1138-1139)
 *   (rho e_tot, rho) = 0                                                          (:1136-1137)
 *   (u_h, u_h) = dtg diag(1 / rho) D - I                                           (:1253-1258)
 * advdiv_matrix and grad_matrix are the centre divergence of a C3 face field
 * with zero boundary flux and the face gradient of a centre field, so row k
 * of D couples centres k - 1, k, k + 1 through faces k and k + 1. */
/* Eddy diffusivity K_h of centre k (ᶜcompute_eddy_diffusivity_coefficient,
 * cache/eddy_diffusivity_coefficient.jl:16-41):
 *   DecayWithHeightDiffusion: D0 exp(-(z - z_sfc) / H)  (precomputed_quantities.jl:651-653),
 *     D0 per column;
 *   VerticalDiffusion: C_E |u_h| z_a with |u_h| the lowest centre's wind speed and
 *     z_a half the lowest layer's depth, times exp(-((p_pbl - p) / p_strato)^2)
 *     where p < p_pbl = 85000 Pa, p_strato = 10000 Pa (precomputed_quantities.jl:671-676).
 * Reads the precomputed p and lowest-centre wind speed. */
static real centre_K(const ora_ctx *c, int k, int col) {
    if (c->n.vdiff_kind == 1) {
        real za = (c->zf[1] - c->zf[0]) / 2;
        real KE = (real)c->n.vdiff_C_E * c->uh0[col] * za;
        real p = c->p[IDX(k, col)];
        if (p > 85000) return KE;
        real x = (85000 - p) / 10000;
        return KE * (real)exp(-(double)(x * x));
    }
    return (real)c->n.D0_col[col] * (real)exp(-(double)(c->zc[k] - c->zf[0]) / c->n.vdiff_H);
}

/* rho K on face f = interp(rho) / interp(1 / max(K, eps))  (vertical_diffusion_boundary_layer.jl:86-90) */
static real face_rhoK(const ora_ctx *c, const real *rho, int f, int col) {
    real eps = (real)(sizeof(real) == 4 ? 1.1920929e-07 : 2.220446049250313e-16);
    real Ka = centre_K(c, f - 1, col), Kb = centre_K(c, f, col);
    Ka = Ka > eps ? Ka : eps;
    Kb = Kb > eps ? Kb : eps;
    return ((rho[IDX(f - 1, col)] + rho[IDX(f, col)]) / 2) / ((1 / Ka + 1 / Kb) / 2);
}

static void diffusion_blocks(ora_ctx *c, real *Y, real dtg, int col) {
    int nl = c->nl;
    real *rho = RHO(Y);
    real cv = (real)c->n.cv_d, cp = (real)c->n.cp_d;
    for (int k = 0; k < nl; ++k) {
        int64_t q = IDX(k, col);
        real Jc = c->Jc[q];
        real dlo = 0, dup = 0;
        if (k >= 1) { /* face k */
            int f = k;
            real rK = face_rhoK(c, rho, f, col);
            real g33 = 1 / (c->Jvf[f] * c->Jvf[f]);
            dlo = c->Jf[IDX(f, col)] * g33 * rK / Jc;
        }
        if (k + 1 <= nl - 1) { /* face k + 1 */
            int f = k + 1;
            real rK = face_rhoK(c, rho, f, col);
            real g33 = 1 / (c->Jvf[f] * c->Jvf[f]);
            dup = c->Jf[IDX(f, col)] * g33 * rK / Jc;
        }
        real dd = -(dlo + dup);
        c->Aee_lo[q] = k >= 1 ? dtg * dlo * (cp / (cv * rho[IDX(k - 1, col)])) : 0;
        c->Aee_d[q] = dtg * dd * (cp / (cv * rho[q])) - 1;
        c->Aee_up[q] = k + 1 <= nl - 1 ? dtg * dup * (cp / (cv * rho[IDX(k + 1, col)])) : 0;
        c->Ahh_lo[q] = dtg * (1 / rho[q]) * dlo;
        c->Ahh_d[q] = dtg * (1 / rho[q]) * dd - 1;
        c->Ahh_up[q] = dtg * (1 / rho[q]) * dup;
    }
}

/* --------------------------------------------------------------- Jacobian */
/* update_advection_jacobian!, manual_sparse_jacobian.jl:713-870, dry, no
 * topography, each block scaled by dtgamma; (u3, u3) includes the -I of the
 * residual.  Fallback diagonals of rho, rho e_tot and u_h are -I (:476-481). */
static void update_jacobian(ora_ctx *c, real *Y, real dtg) {
    int nl = c->nl, nc = c->ncol;
    real *rho = RHO(Y), *u3 = U3(Y);
    real Rd = (real)c->n.R_d, cp = (real)c->n.cp_d, cv = (real)c->n.cv_d, T0 = (real)c->n.T_0;
    real kap = Rd / cv;
    const double *G = c->g.gup;
    c->dtgamma = dtg;
    real *Pi = c->t[0], *thv = c->t[2], *dpdr = c->t[3];
    for (int k = 0; k < nl; ++k)
        ORA_PAR_FOR
        for (int col = 0; col < nc; ++col) {
            int64_t q = IDX(k, col);
            real P = exner(c, c->p[q]);
            Pi[q] = P;
            thv[q] = c->T[q] / P;
            dpdr[q] = kap * (T0 * cp - c->K[q] - c->Phic[k]) + (Rd - kap * cv) * c->T[q];
        }
    ORA_PAR_FOR
    for (int col = 0; col < nc; ++col) {
        /* centre rows: A(rho, u3) = dtg * adv * diag(g33), adv = -advdiv_matrix * diag(I^f(rho J)/J_f) */
        for (int k = 0; k < nl; ++k) {
            int64_t q = IDX(k, col);
            real Jc = c->Jc[q];
            real a0 = 0, a1 = 0, e0 = 0, e1 = 0;
            if (k >= 1) {
                int f = k;
                real rJ = (rho[IDX(f - 1, col)] * c->Jc[IDX(f - 1, col)] + rho[q] * Jc) / 2;
                real Jf = c->Jf[IDX(f, col)];
                real g33 = 1 / (c->Jvf[f] * c->Jvf[f]);
                real adv = (Jf * (rJ / Jf)) / Jc;
                real hf = (c->h[IDX(f - 1, col)] + c->h[q]) / 2;
                a0 = dtg * adv * g33;
                e0 = dtg * adv * (hf * g33);
            }
            if (k + 1 <= nl - 1) {
                int f = k + 1;
                real rJ = (rho[q] * Jc + rho[IDX(f, col)] * c->Jc[IDX(f, col)]) / 2;
                real Jf = c->Jf[IDX(f, col)];
                real g33 = 1 / (c->Jvf[f] * c->Jvf[f]);
                real adv = -(Jf * (rJ / Jf)) / Jc;
                real hf = (c->h[q] + c->h[IDX(f, col)]) / 2;
                a1 = dtg * adv * g33;
                e1 = dtg * adv * (hf * g33);
            }
            c->Aru0[q] = a0;
            c->Aru1[q] = a1;
            c->Aeu0[q] = e0;
            c->Aeu1[q] = e1;
        }
        /* face rows */
        for (int f = 0; f <= nl; ++f) {
            int64_t q = IDX(f, col);
            real beta = c->n.rayleigh ? alpha_w_of(c, col) * rayleigh_zeta(c, c->zf[f]) : 0;
            if (f == 0 || f == nl) {
                c->Aur0[q] = c->Aur1[q] = c->Aue0[q] = c->Aue1[q] = 0;
                c->Auh0a[q] = c->Auh0b[q] = c->Auh1a[q] = c->Auh1b[q] = 0;
                c->Auu_lo[q] = c->Auu_up[q] = 0;
                c->Auu_d[q] = dtg * (-beta) - 1;
                continue;
            }
            int64_t a = IDX(f - 1, col), b = IDX(f, col);
            real rf = (rho[a] + rho[b]) / 2;
            /* p_grad_matrix = diag(-1 / I^f(rho)) * grad_matrix: (-1, +1) on (c=f-1, c=f) */
            real pg_m = -(-1 / rf), pg_p = -1 / rf;  /* entries for centre f-1 and centre f */
            real buoy = cp * ((thv[a] + thv[b]) / 2) * (Pi[b] - Pi[a]) / rf;
            c->Aur0[q] = dtg * (pg_m * dpdr[a] + buoy / 2);
            c->Aur1[q] = dtg * (pg_p * dpdr[b] + buoy / 2);
            c->Aue0[q] = dtg * (pg_m * kap);
            c->Aue1[q] = dtg * (pg_p * kap);
            /* d K / d u_h = CT12(u_h) (:746-751) */
            real sa = c->sc[f - 1], sb = c->sc[f];
            real ua1, ua2, ub1, ub2;
            {
                real iv = 1 / (sa * sa);
                int colc = col;
                real x1 = UH1(Y)[a], x2 = UH2(Y)[a];
                ua1 = iv * ((real)G[colc] * x1 + (real)G[nc + colc] * x2);
                ua2 = iv * ((real)G[nc + colc] * x1 + (real)G[2 * nc + colc] * x2);
                iv = 1 / (sb * sb);
                x1 = UH1(Y)[b];
                x2 = UH2(Y)[b];
                ub1 = iv * ((real)G[colc] * x1 + (real)G[nc + colc] * x2);
                ub2 = iv * ((real)G[nc + colc] * x1 + (real)G[2 * nc + colc] * x2);
            }
            real ka = -kap * rho[a], kb = -kap * rho[b];
            c->Auh0a[q] = dtg * pg_m * ka * ua1;
            c->Auh0b[q] = dtg * pg_m * ka * ua2;
            c->Auh1a[q] = dtg * pg_p * kb * ub1;
            c->Auh1b[q] = dtg * pg_p * kb * ub2;
            /* d K_c / d u3_f = 1/2 CT3(u3)_f for f in {c, c+1} (:752-754) */
            real gm = 1 / (c->Jvf[f - 1] * c->Jvf[f - 1]), g0 = 1 / (c->Jvf[f] * c->Jvf[f]),
                 gp = 1 / (c->Jvf[f + 1] * c->Jvf[f + 1]);
            real dKa_m = u3[IDX(f - 1, col)] * gm / 2, dKa_0 = u3[q] * g0 / 2; /* centre f-1 */
            real dKb_0 = u3[q] * g0 / 2, dKb_p = u3[IDX(f + 1, col)] * gp / 2; /* centre f */
            c->Auu_lo[q] = dtg * (pg_m * ka * dKa_m);
            c->Auu_d[q] = dtg * (pg_m * ka * dKa_0 + pg_p * kb * dKb_0 + (-beta)) - 1;
            c->Auu_up[q] = dtg * (pg_p * kb * dKb_p);
        }
        if (c->n.vert_diff && c->n.implicit_diffusion) diffusion_blocks(c, Y, dtg, col);
    }
}
void ora_update_jacobian(ora_ctx *c, ora_real *Y, double dtgamma) { update_jacobian(c, (real *)Y, (real)dtgamma); }

/* ldiv!: BlockArrowheadSolve(rho, rho e_tot; alg2 = BlockLowerTriangularSolve(u_h))
 * (manual_sparse_jacobian.jl:579-584).  Schur complement on u3:
 *   S = A(u3,u3) + A(u3,rho) A(rho,u3) + A(u3,rhoe) A(rhoe,u3)  (tridiagonal)
 *   S du3 = R3 + A(u3,rho) Rrho + A(u3,rhoe) Rrhoe + A(u3,uh) Ruh,
 *   duh = -Ruh, drho = A(rho,u3) du3 - Rrho, drhoe = A(rhoe,u3) du3 - Rrhoe. */
static void jac_ldiv(ora_ctx *c, real *dY, const real *R) {
    int nl = c->nl, nc = c->ncol;
    real *lo = c->t[4], *di = c->t[5], *up = c->t[6], *rhs = c->t[7], *cp_ = c->t[8];
    ORA_PAR_FOR
    for (int col = 0; col < nc; ++col) {
        for (int f = 0; f <= nl; ++f) {
            int64_t q = IDX(f, col);
            real L = c->Auu_lo[q], Dg = c->Auu_d[q], U = c->Auu_up[q];
            real r = R[4 * c->nC + q];
            if (f >= 1 && f <= nl - 1) {
                int64_t a = IDX(f - 1, col), b = IDX(f, col);
                /* A(u3,rho)[f][f-1] * A(rho,u3)[f-1][f-1 or f], and [f][f] * A(rho,u3)[f][f or f+1] */
                L += c->Aur0[q] * c->Aru0[a] + c->Aue0[q] * c->Aeu0[a];
                Dg += c->Aur0[q] * c->Aru1[a] + c->Aue0[q] * c->Aeu1[a] + c->Aur1[q] * c->Aru0[b] +
                      c->Aue1[q] * c->Aeu0[b];
                U += c->Aur1[q] * c->Aru1[b] + c->Aue1[q] * c->Aeu1[b];
                r += c->Aur0[q] * R[a] + c->Aur1[q] * R[b] + c->Aue0[q] * R[3 * c->nC + a] +
                     c->Aue1[q] * R[3 * c->nC + b];
                r += c->Auh0a[q] * R[c->nC + a] + c->Auh0b[q] * R[2 * c->nC + a] +
                     c->Auh1a[q] * R[c->nC + b] + c->Auh1b[q] * R[2 * c->nC + b];
            }
            lo[q] = L;
            di[q] = Dg;
            up[q] = U;
            rhs[q] = r;
        }
        /* Thomas algorithm */
        cp_[IDX(0, col)] = up[IDX(0, col)] / di[IDX(0, col)];
        rhs[IDX(0, col)] = rhs[IDX(0, col)] / di[IDX(0, col)];
        for (int f = 1; f <= nl; ++f) {
            int64_t q = IDX(f, col), pq = IDX(f - 1, col);
            real m = di[q] - lo[q] * cp_[pq];
            cp_[q] = up[q] / m;
            rhs[q] = (rhs[q] - lo[q] * rhs[pq]) / m;
        }
        real *du3 = dY + 4 * c->nC;
        du3[IDX(nl, col)] = rhs[IDX(nl, col)];
        for (int f = nl - 1; f >= 0; --f) du3[IDX(f, col)] = rhs[IDX(f, col)] - cp_[IDX(f, col)] * du3[IDX(f + 1, col)];
        for (int k = 0; k < nl; ++k) {
            int64_t q = IDX(k, col);
            real x0 = du3[IDX(k, col)], x1 = du3[IDX(k + 1, col)];
            dY[q] = c->Aru0[q] * x0 + c->Aru1[q] * x1 - R[q];
            dY[3 * c->nC + q] = c->Aeu0[q] * x0 + c->Aeu1[q] * x1 - R[3 * c->nC + q];
            dY[c->nC + q] = -R[c->nC + q];
            dY[2 * c->nC + q] = -R[2 * c->nC + q];
        }
    }
}

/* Thomas algorithm for one tridiagonal system of n rows (lo[0], up[n-1] unused) */
static void tridiag(int n, const real *lo, const real *d, const real *up, const real *r, real *x, real *w) {
    real m = d[0];
    x[0] = r[0] / m;
    for (int i = 1; i < n; ++i) {
        w[i] = up[i - 1] / m;
        m = d[i] - lo[i] * w[i];
        x[i] = (r[i] - lo[i] * x[i - 1]) / m;
    }
    for (int i = n - 2; i >= 0; --i) x[i] -= w[i + 1] * x[i + 1];
}

/* ldiv! with implicit diffusion: ApproximateBlockArrowheadIterativeSolve(rho,
 * rho e_tot; alg1 = BlockLowerTriangularSolve(rho), alg2 =
 * BlockLowerTriangularSolve(u_h), P_alg1 = MainDiagonalPreconditioner(),
 * n_iters) (manual_sparse_jacobian.jl:538-578), restated from ClimaCore
 * 0.15.1's MatrixFields (not vendored): a Schur complement reduction on the
 * scalars (rho, rho e_tot) whose scalar block A11 is inverted exactly, with
 * the velocity Schur complement S = A22 - A21 inv(A11) A12 solved by
 * n_iters stationary iterations x <- x - inv(P) (S x - b2') from x = 0,
 * preconditioned by P = A22 - A21 inv(diag(A11)) A12 (tridiagonal on u_3),
 * which is solved block lower triangularly (u_h, then u_3).  Without
 * diffusion diag(A11) = A11 = -I and one iteration is the exact
 * BlockArrowheadSolve of jac_ldiv. */
static void jac_ldiv_approx(ora_ctx *c, real *dY, const real *R) {
    const int nl = c->nl, nc = c->ncol, nf = nl + 1;
    const int64_t nC = c->nC;
    const int iters = c->n.solve_iters > 0 ? c->n.solve_iters : 1;
    ORA_PAR_FOR
    for (int col = 0; col < nc; ++col) {
        real elo[nl], ed[nl], eup[nl], hlo[nl], hd[nl], hup[nl];
        real br[nl], bh1[nl], bh2[nl], be[nl], bw[nf];
        real ye[nl], xh1[nl], xh2[nl], xw[nf], rh1[nl], rh2[nl], rw[nf], zh1[nl], zh2[nl], zw[nf];
        real slo[nf], sd[nf], sup[nf], g[nl], m[nl], wk[nf > nl ? nf : nl];
#define CB(a, k) (c->a[IDX(k, col)])
        for (int k = 0; k < nl; ++k) {
            elo[k] = CB(Aee_lo, k); ed[k] = CB(Aee_d, k); eup[k] = CB(Aee_up, k);
            hlo[k] = CB(Ahh_lo, k); hd[k] = CB(Ahh_d, k); hup[k] = CB(Ahh_up, k);
            br[k] = R[IDX(k, col)];
            bh1[k] = R[nC + IDX(k, col)];
            bh2[k] = R[2 * nC + IDX(k, col)];
            be[k] = R[3 * nC + IDX(k, col)];
        }
        for (int f = 0; f <= nl; ++f) bw[f] = R[4 * nC + IDX(f, col)];
        /* y1 = inv(A11) b1: y_rho = -b_rho, A_ee y_e = b_e */
        tridiag(nl, elo, ed, eup, be, ye, wk);
        /* b2' = b2 - A21 y1 (u_3 rows only) */
        for (int f = 1; f < nl; ++f)
            bw[f] -= CB(Aur0, f) * (-br[f - 1]) + CB(Aur1, f) * (-br[f]) + CB(Aue0, f) * ye[f - 1] + CB(Aue1, f) * ye[f];
        /* the preconditioner's u_3 block P_ww = A_ww - A_wrho (-1) A_rhow - A_we diag(A_ee)^-1 A_ew */
        for (int f = 0; f <= nl; ++f) {
            slo[f] = CB(Auu_lo, f);
            sd[f] = CB(Auu_d, f);
            sup[f] = CB(Auu_up, f);
            if (f >= 1 && f <= nl - 1) {
                int a = f - 1, b = f;
                slo[f] += CB(Aur0, f) * CB(Aru0, a) - CB(Aue0, f) * CB(Aeu0, a) / ed[a];
                sd[f] += CB(Aur0, f) * CB(Aru1, a) + CB(Aur1, f) * CB(Aru0, b) -
                         (CB(Aue0, f) * CB(Aeu1, a) / ed[a] + CB(Aue1, f) * CB(Aeu0, b) / ed[b]);
                sup[f] += CB(Aur1, f) * CB(Aru1, b) - CB(Aue1, f) * CB(Aeu1, b) / ed[b];
            }
        }
        for (int k = 0; k < nl; ++k) xh1[k] = xh2[k] = 0;
        for (int f = 0; f <= nl; ++f) xw[f] = 0;
        for (int it = 0; it < iters; ++it) {
            /* r = b2' - S x, S x = A22 x + A_wrho A_rhow x_w - A_we inv(A_ee) A_ew x_w */
            for (int k = 0; k < nl; ++k) {
                real s1 = hd[k] * xh1[k], s2 = hd[k] * xh2[k];
                if (k >= 1) { s1 += hlo[k] * xh1[k - 1]; s2 += hlo[k] * xh2[k - 1]; }
                if (k + 1 < nl) { s1 += hup[k] * xh1[k + 1]; s2 += hup[k] * xh2[k + 1]; }
                rh1[k] = bh1[k] - s1;
                rh2[k] = bh2[k] - s2;
                g[k] = CB(Aeu0, k) * xw[k] + CB(Aeu1, k) * xw[k + 1];
            }
            tridiag(nl, elo, ed, eup, g, m, wk);
            for (int f = 0; f <= nl; ++f) {
                real sx = CB(Auu_d, f) * xw[f];
                if (f >= 1) sx += CB(Auu_lo, f) * xw[f - 1];
                if (f + 1 <= nl) sx += CB(Auu_up, f) * xw[f + 1];
                if (f >= 1 && f <= nl - 1) {
                    int a = f - 1, b = f;
                    sx += CB(Auh0a, f) * xh1[a] + CB(Auh0b, f) * xh2[a] + CB(Auh1a, f) * xh1[b] + CB(Auh1b, f) * xh2[b];
                    real pa = CB(Aru0, a) * xw[a] + CB(Aru1, a) * xw[a + 1];
                    real pb = CB(Aru0, b) * xw[b] + CB(Aru1, b) * xw[b + 1];
                    sx += CB(Aur0, f) * pa + CB(Aur1, f) * pb;
                    sx -= CB(Aue0, f) * m[a] + CB(Aue1, f) * m[b];
                }
                rw[f] = bw[f] - sx;
            }
            /* z = inv(P) r: u_h first, then u_3 with the (u_3, u_h) coupling */
            tridiag(nl, hlo, hd, hup, rh1, zh1, wk);
            tridiag(nl, hlo, hd, hup, rh2, zh2, wk);
            for (int f = 1; f < nl; ++f)
                rw[f] -= CB(Auh0a, f) * zh1[f - 1] + CB(Auh0b, f) * zh2[f - 1] + CB(Auh1a, f) * zh1[f] +
                         CB(Auh1b, f) * zh2[f];
            tridiag(nf, slo, sd, sup, rw, zw, wk);
            for (int k = 0; k < nl; ++k) { xh1[k] += zh1[k]; xh2[k] += zh2[k]; }
            for (int f = 0; f <= nl; ++f) xw[f] += zw[f];
        }
        /* x1 = inv(A11) (b1 - A12 x2) */
        for (int k = 0; k < nl; ++k) g[k] = be[k] - (CB(Aeu0, k) * xw[k] + CB(Aeu1, k) * xw[k + 1]);
        tridiag(nl, elo, ed, eup, g, m, wk);
        for (int k = 0; k < nl; ++k) {
            int64_t q = IDX(k, col);
            dY[q] = CB(Aru0, k) * xw[k] + CB(Aru1, k) * xw[k + 1] - br[k];
            dY[nC + q] = xh1[k];
            dY[2 * nC + q] = xh2[k];
            dY[3 * nC + q] = m[k];
        }
        for (int f = 0; f <= nl; ++f) dY[4 * nC + IDX(f, col)] = xw[f];
#undef CB
    }
}
void ora_ldiv(ora_ctx *c, ora_real *dY, const ora_real *R) {
    if (c->n.vert_diff && c->n.implicit_diffusion)
        jac_ldiv_approx(c, (real *)dY, (const real *)R);
    else
        jac_ldiv(c, (real *)dY, (const real *)R);
}

void ora_column_blocks(ora_ctx *c, int col, double *Aru, double *Aeu, double *Aur, double *Aue,
                       double *Auu_lo, double *Auu_d, double *Auu_up) {
    for (int k = 0; k < c->nl; ++k) {
        Aru[2 * k] = c->Aru0[IDX(k, col)];
        Aru[2 * k + 1] = c->Aru1[IDX(k, col)];
        Aeu[2 * k] = c->Aeu0[IDX(k, col)];
        Aeu[2 * k + 1] = c->Aeu1[IDX(k, col)];
    }
    for (int f = 0; f <= c->nl; ++f) {
        Aur[2 * f] = c->Aur0[IDX(f, col)];
        Aur[2 * f + 1] = c->Aur1[IDX(f, col)];
        Aue[2 * f] = c->Aue0[IDX(f, col)];
        Aue[2 * f + 1] = c->Aue1[IDX(f, col)];
        Auu_lo[f] = c->Auu_lo[IDX(f, col)];
        Auu_d[f] = c->Auu_d[IDX(f, col)];
        Auu_up[f] = c->Auu_up[IDX(f, col)];
    }
}

void ora_column_diffusion_blocks(ora_ctx *c, int col, double *Aee, double *Ahh) {
    for (int k = 0; k < c->nl; ++k) {
        Aee[3 * k] = c->Aee_lo[IDX(k, col)];
        Aee[3 * k + 1] = c->Aee_d[IDX(k, col)];
        Aee[3 * k + 2] = c->Aee_up[IDX(k, col)];
        Ahh[3 * k] = c->Ahh_lo[IDX(k, col)];
        Ahh[3 * k + 1] = c->Ahh_d[IDX(k, col)];
        Ahh[3 * k + 2] = c->Ahh_up[IDX(k, col)];
    }
}

void ora_column_uh_blocks(ora_ctx *c, int col, double *Auh) {
    for (int f = 0; f <= c->nl; ++f) {
        Auh[4 * f] = c->Auh0a[IDX(f, col)];
        Auh[4 * f + 1] = c->Auh0b[IDX(f, col)];
        Auh[4 * f + 2] = c->Auh1a[IDX(f, col)];
        Auh[4 * f + 3] = c->Auh1b[IDX(f, col)];
    }
}

/* correct_implicit_advection_tendency!, implicit_tendency.jl:322-339:
 * Yt.rhoe = vtt(upwind) - vtt(central), all other components zero. */
static void correct_implicit(ora_ctx *c, real *Yt, real *Y) {
    memset(Yt, 0, sizeof(real) * c->S);
    if (c->n.upwinding == 0) return;
    real *a = c->t[8], *b = c->t[9];
    vertical_transport(c, RHO(Y), c->h, c->n.upwinding, a);
    vertical_transport(c, RHO(Y), c->h, 0, b);
    ORA_PAR_FOR
    for (int64_t q = 0; q < c->nC; ++q) RHOE(Yt)[q] = a[q] - b[q];
}
void ora_correct_implicit(ora_ctx *c, ora_real *Yt, ora_real *Y) { correct_implicit(c, (real *)Yt, (real *)Y); }

/* ------------------------------------------------------ explicit tendency */

/* one level of C123(wgrad(div X)) and C123(wcurl(C123(curl X))) for a C123
 * field X = (x1, x2, x3) at centre level k; outputs gd (C12) and wc (C123). */
static void vector_laplacian_parts(ora_ctx *c, int k, const real *x1, const real *x2, const real *x3,
                                   real *gd1, real *gd2, real *wc1, real *wc2, real *wc3) {
    int nc = c->ncol;
    real *J = c->Jc + (int64_t)k * nc;
    real s = c->sc[k], Jv = c->Jvc[k];
    real *v1 = c->t[0] + 0, *v2 = c->t[0] + nc, *dv = c->t[1];
    real *cr1 = c->t[2], *cr2 = c->t[2] + nc, *cr3 = c->t[3];
    real *cc1 = c->t[4], *cc2 = c->t[4] + nc, *cc3 = c->t[5];
    real *w1 = c->t[6], *w2 = c->t[6] + nc, *w3 = c->t[7];
    ORA_PAR_FOR
    for (int col = 0; col < nc; ++col) to_contra(c, col, s, x1[col], x2[col], &v1[col], &v2[col]);
    op_div(c, J, v1, v2, dv);
    op_wgrad(c, dv, gd1, gd2);
    op_curl3(c, J, x1, x2, cr3);
    op_curl12(c, J, x3, cr1, cr2);
    ORA_PAR_FOR
    for (int col = 0; col < nc; ++col) {
        to_cov(c, col, s, cr1[col], cr2[col], &cc1[col], &cc2[col]);
        cc3[col] = Jv * Jv * cr3[col];
    }
    op_wcurl3(c, J, cc1, cc2, w3);
    op_wcurl12(c, J, cc3, w1, w2);
    ORA_PAR_FOR
    for (int col = 0; col < nc; ++col) {
        to_cov(c, col, s, w1[col], w2[col], &wc1[col], &wc2[col]);
        wc3[col] = Jv * Jv * w3[col];
    }
}

/* remaining_tendency!, remaining_tendency.jl:48-58; precomputed quantities
 * must have been set from Y (the time stepper's cache! hook). */
/* ----------------------------------------------------- surface fluxes */
/* surface_flux_tendency!, surface_flux.jl:101-145, with sfc_conditions from
 * update_surface_conditions! (surface_conditions.jl:22-70) for the default
 * AtmosSurface (types.jl:2292-2302): ExchangeCoefficients(Cd, Ch) with an
 * AnalyticTemperature surface temperature, no boundary overrides (surface
 * winds 0, gustiness from the config), dry air.  Called from
 * additional_tendency!, i.e. explicit; uses the precomputed T of the cache.
 *
 * The flux closure itself is SurfaceFluxes 1.2.1 (not vendored; Manifest
 * git-tree-sha1 0b35705e...), restated from its published bulk formulas:
 *   wind speed   U = max(|u_int - u_sfc|, gustiness)
 *   momentum     rho tau_xz = -rho_sfc Cd U (u_int - u_sfc), same for y
 *   heat         shf = -rho_sfc Ch U (DSE_int - DSE_sfc),
 *                DSE = cp_d (T - T_0) + Phi (dry static energy)
 *   density      rho_sfc: the interior pressure extrapolated hydrostatically
 *                over dz = z_int - z_sfc at the layer's mean temperature,
 *                p_sfc = p_int exp(g dz / (R_d (T_int + T_sfc) / 2)),
 *                rho_sfc = p_sfc / (R_d T_sfc)
 * Dry air: lhf = evaporation = 0.  The interior wind components are
 * projected_vector_data(CT1 / CT2, u, L) (utilities.jl:526-527): the
 * contravariant component times |a_i| = sqrt(g_ii); the flux tensor is
 * n3 (x) C12(rho tau_xz xz + rho tau_yz yz) with xz, yz the CT1, CT2 unit
 * vectors at the surface and n3 = C3(unit_basis_vector_data(C3, L))
 * (tensor_from_components, surface_conditions.jl:357-362).  The tendencies
 * are DivergenceF2C with bottom SetValue(flux) and a zero interior flux
 * (boundary_tendency_scalar / _momentum, surface_flux.jl:35-75): only the
 * first centre changes, by +J_f F^3 / J_c (divided by rho for u_h). */
static void surface_flux_tendency(ora_ctx *c, real *Yt, real *Y) {
    const int nc = c->ncol;
    const double *G = c->g.gup, *Gd = c->g.gdn;
    const real *rho = RHO(Y), *u1 = UH1(Y), *u2 = UH2(Y);
    const real R = (real)c->n.R_d, cpd = (real)c->n.cp_d, T0 = (real)c->n.T_0, grav = (real)c->n.grav;
    const real Cd = (real)c->n.Cd, Ch = (real)c->n.Ch, gust = (real)c->n.gustiness;
    ORA_PAR_FOR
    for (int col = 0; col < nc; ++col) {
        const int64_t q = IDX(0, col), qf = IDX(0, col);
        const real rho_int = rho[q], T_int = c->T[q];
        const real z_int = c->zc[0], z_sfc = c->zf[0];
        const real dz = z_int - z_sfc;
        const real Phi_int = grav * z_int, Phi_sfc = grav * z_sfc;
        const real T_sfc = (real)c->n.T_sfc_col[col];
        /* interior velocity CT1, CT2 (no topography: g^13 = g^23 = 0) */
        const real s = c->sc[0], is2 = 1 / (s * s);
        const real ct1 = is2 * ((real)G[col] * u1[q] + (real)G[nc + col] * u2[q]);
        const real ct2 = is2 * ((real)G[nc + col] * u1[q] + (real)G[2 * nc + col] * u2[q]);
        const real u_int = ct1 * sqrt((real)Gd[col] * s * s);
        const real v_int = ct2 * sqrt((real)Gd[2 * nc + col] * s * s);
        /* SurfaceFluxes: bulk fluxes with Cd, Ch */
        real speed = sqrt(u_int * u_int + v_int * v_int);
        if (speed < gust) speed = gust;
        const real p_int = rho_int * R * T_int;
        const real p_sfc = p_int * exp(grav * dz / (R * (T_int + T_sfc) / 2));
        const real rho_sfc = p_sfc / (R * T_sfc);
        const real dse_int = cpd * (T_int - T0) + Phi_int, dse_sfc = cpd * (T_sfc - T0) + Phi_sfc;
        const real shf = -rho_sfc * Ch * speed * (dse_int - dse_sfc);
        const real tau_x = -rho_sfc * Cd * speed * u_int, tau_y = -rho_sfc * Cd * speed * v_int;
        /* flux vectors on the surface normal, with the surface metric */
        const real s0 = c->sf[0];
        const real g11 = (real)Gd[col] * s0 * s0, g12 = (real)Gd[nc + col] * s0 * s0,
                   g22 = (real)Gd[2 * nc + col] * s0 * s0;
        const real fc1 = tau_x / sqrt(g11), fc2 = tau_y / sqrt(g22); /* CT12 */
        const real f1 = g11 * fc1 + g12 * fc2, f2 = g12 * fc1 + g22 * fc2; /* C12 */
        const real g33 = 1 / (c->Jvf[0] * c->Jvf[0]);
        const real n3 = 1 / sqrt(g33); /* C3 component of the unit normal */
        /* DivergenceF2C at the first centre: (J F^3)(face 1, zero) - (J F^3)(face 0) */
        const real Jf0 = c->Jf[qf], Jc0 = c->Jc[q];
        const real div_e = (0 - Jf0 * (g33 * (shf * n3))) / Jc0;
        const real div_1 = (0 - Jf0 * (g33 * (n3 * f1))) / Jc0;
        const real div_2 = (0 - Jf0 * (g33 * (n3 * f2))) / Jc0;
        RHOE(Yt)[q] -= div_e;
        UH1(Yt)[q] -= div_1 / rho_int;
        UH2(Yt)[q] -= div_2 / rho_int;
    }
}

/* ------------------------------------------------- vertical diffusion */
/* vertical_diffusion_boundary_layer_tendency!, vertical_diffusion_boundary_layer.jl:63-103,
 * with K_h from centre_K (DecayWithHeightDiffusion or VerticalDiffusion).  Added to Yt: by
 * the explicit tendency (additional_tendency!, remaining_tendency.jl:185-190)
 * with diff_mode Explicit, by the implicit tendency (implicit_tendency.jl:
 * 73-82) with diff_mode Implicit.  Reads the precomputed T. */
static void vertical_diffusion_tendency(ora_ctx *c, real *Yt, real *Y) {
    int nl = c->nl, nc = c->ncol;
    int sphere = c->g.nelem > 0;
    real *rho = RHO(Y), *u1 = UH1(Y), *u2 = UH2(Y);
    real cpd = (real)c->n.cp_d, T0 = (real)c->n.T_0;
    const double *B = c->g.basis;
    ORA_PAR_FOR
    for (int col = 0; col < nc; ++col) {
        /* s_d flux: -rhoK grad_v(s_d); diffdiv with zero boundary flux */
        real *F = c->t[10], *Fu1 = c->t[0], *Fu2 = c->t[1], *Fu3 = c->t[2];
        for (int f = 0; f <= nl; ++f) {
            int64_t q = IDX(f, col);
            F[q] = 0;
            Fu1[q] = Fu2[q] = Fu3[q] = 0;
            if (f == 0 || f == nl) continue;
            int64_t a = IDX(f - 1, col), b = IDX(f, col);
            real rK = face_rhoK(c, rho, f, col);
            real sda = cpd * (c->T[a] - T0) + c->Phic[f - 1];
            real sdb = cpd * (c->T[b] - T0) + c->Phic[f];
            real g33 = 1 / (c->Jvf[f] * c->Jvf[f]);
            F[q] = c->Jf[q] * (g33 * (-(rK * (sdb - sda))));
            /* momentum: C3 (x) UVW strain; flux row of -2 rhoK S */
            real ua[3], ub[3];
            if (sphere) {
                for (int d = 0; d < 3; ++d) {
                    ua[d] = (u1[a] * (real)B[d * nc + col] + u2[a] * (real)B[(3 + d) * nc + col]) / c->sc[f - 1];
                    ub[d] = (u1[b] * (real)B[d * nc + col] + u2[b] * (real)B[(3 + d) * nc + col]) / c->sc[f];
                }
            } else {
                ua[0] = u1[a]; ua[1] = u2[a]; ua[2] = 0;
                ub[0] = u1[b]; ub[1] = u2[b]; ub[2] = 0;
            }
            real inv = 1 / c->Jvf[f];
            real fac = c->Jf[q] * inv * (-2 * rK) / 2;
            Fu1[q] = fac * (ub[0] - ua[0]) * inv;
            Fu2[q] = fac * (ub[1] - ua[1]) * inv;
            Fu3[q] = fac * (ub[2] - ua[2]) * inv;
        }
        for (int k = 0; k < nl; ++k) {
            int64_t q = IDX(k, col);
            real Jc = c->Jc[q];
            RHOE(Yt)[q] -= (F[IDX(k + 1, col)] - F[IDX(k, col)]) / Jc;
            real d[3] = {(Fu1[IDX(k + 1, col)] - Fu1[IDX(k, col)]) / Jc,
                         (Fu2[IDX(k + 1, col)] - Fu2[IDX(k, col)]) / Jc,
                         (Fu3[IDX(k + 1, col)] - Fu3[IDX(k, col)]) / Jc};
            real c1, c2;
            if (sphere) {
                c1 = c2 = 0;
                for (int e = 0; e < 3; ++e) {
                    c1 += d[e] * (real)B[(6 + e) * nc + col];
                    c2 += d[e] * (real)B[(9 + e) * nc + col];
                }
                c1 *= c->sc[k];
                c2 *= c->sc[k];
            } else {
                c1 = d[0];
                c2 = d[1];
            }
            UH1(Yt)[q] -= c1 / rho[q];
            UH2(Yt)[q] -= c2 / rho[q];
        }
    }
}

static void remaining_tendency(ora_ctx *c, real *Yt, real *Y) {
    int nl = c->nl, nc = c->ncol;
    int sphere = c->g.nelem > 0;
    memset(Yt, 0, sizeof(real) * c->S);
    real cp = (real)c->n.cp_d;
    real *rho = RHO(Y), *u1 = UH1(Y), *u2 = UH2(Y), *u3 = U3(Y);
    real *e4 = c->t[8]; /* nabla^4 u_3 at centres (for the face interpolation) */
    real *om3 = c->t[9]; /* centre omega^3 */
    if (sphere) {
        /* ---- horizontal_dynamics_tendency!, advection.jl:36-91 ---- */
        for (int k = 0; k < nl; ++k) {
            int64_t o = (int64_t)k * nc;
            real *J = c->Jc + o;
            real s = c->sc[k];
            real *m1 = c->t[0], *m2 = c->t[0] + nc, *wd = c->t[1], *wdh = c->t[1] + nc;
            real *mh1 = c->t[2], *mh2 = c->t[2] + nc, *gh1 = c->t[3], *gh2 = c->t[3] + nc;
            ORA_PAR_FOR
            for (int col = 0; col < nc; ++col) {
                real c1, c2;
                to_contra(c, col, s, u1[o + col], u2[o + col], &c1, &c2);
                m1[col] = rho[o + col] * c1;
                m2[col] = rho[o + col] * c2;
                mh1[col] = m1[col] * c->h[o + col];
                mh2[col] = m2[col] * c->h[o + col];
            }
            op_wdiv(c, J, m1, m2, wd);
            /* split_div(rho u, 1) = 1/2 wdiv(rho u) + 1/2 (wdiv(rho u) + 0) */
            ORA_PAR_FOR
            for (int col = 0; col < nc; ++col) RHO(Yt)[o + col] -= wd[col] / 2 + (wd[col] + 0) / 2;
            op_wdiv(c, J, mh1, mh2, wdh);
            op_grad(c, c->h + o, gh1, gh2);
            ORA_PAR_FOR
            for (int col = 0; col < nc; ++col) {
                real sd = wdh[col] / 2 + (c->h[o + col] * wd[col] + (m1[col] * gh1[col] + m2[col] * gh2[col])) / 2;
                RHOE(Yt)[o + col] -= sd;
            }
            /* u_h -= C12(grad(K + Phi - Phi_r) + cp_d (th' grad Pi + grad(th' Pi) - Pi grad th') / 2) */
            real *A = c->t[4], *Pi = c->t[4] + nc, *th = c->t[5], *thPi = c->t[5] + nc;
            ORA_PAR_FOR
            for (int col = 0; col < nc; ++col) {
                real P = exner(c, c->p[o + col]);
                Pi[col] = P;
                A[col] = c->K[o + col] + c->Phic[k] - Phi_ref(c, P);
                th[col] = c->T[o + col] / P - T_ref(c, P) / P;
                thPi[col] = th[col] * P;
            }
            real *g1 = c->t[6], *g2 = c->t[6] + nc, *p1 = c->t[7], *p2 = c->t[7] + nc;
            real *q1 = c->t[10], *q2 = c->t[10] + nc, *r1 = c->t[11], *r2 = c->t[11] + nc;
            op_grad(c, A, g1, g2);
            op_grad(c, Pi, p1, p2);
            op_grad(c, thPi, q1, q2);
            op_grad(c, th, r1, r2);
            ORA_PAR_FOR
            for (int col = 0; col < nc; ++col) {
                UH1(Yt)[o + col] -= g1[col] + cp * (th[col] * p1[col] + q1[col] - Pi[col] * r1[col]) / 2;
                UH2(Yt)[o + col] -= g2[col] + cp * (th[col] * p2[col] + q2[col] - Pi[col] * r2[col]) / 2;
            }
        }
        /* ---- hyperdiffusion_tendency!, remaining_tendency.jl:15-24 ---- */
        if (c->n.hyperdiff) {
            /* prep: hyperdiffusion.jl:141-147 */
            for (int k = 0; k < nl; ++k) {
                int64_t o = (int64_t)k * nc;
                real *gd1 = c->t[8], *gd2 = c->t[8] + nc, *wc1 = c->t[9], *wc2 = c->t[9] + nc, *wc3 = c->t[10];
                vector_laplacian_parts(c, k, u1 + o, u2 + o, c->cu3 + o, gd1, gd2, wc1, wc2, wc3);
                ORA_PAR_FOR
                for (int col = 0; col < nc; ++col) {
                    c->L1[o + col] = gd1[col] - wc1[col];
                    c->L2[o + col] = gd2[col] - wc2[col];
                    c->L3[o + col] = 0 - wc3[col];
                }
                real *sd = c->t[11], *q1 = c->t[0], *q2 = c->t[0] + nc, *v1 = c->t[1], *v2 = c->t[1] + nc;
                real cpd = (real)c->n.cp_d, T0 = (real)c->n.T_0;
                ORA_PAR_FOR
                for (int col = 0; col < nc; ++col) {
                    real P = exner(c, c->p[o + col]);
                    real sdr = cpd * (T_ref(c, P) - T0) + Phi_ref(c, P);
                    sd[col] = (cpd * (c->T[o + col] - T0) + c->Phic[k]) - sdr;
                }
                op_grad(c, sd, q1, q2);
                ORA_PAR_FOR
                for (int col = 0; col < nc; ++col) to_contra(c, col, c->sc[k], q1[col], q2[col], &v1[col], &v2[col]);
                op_wdiv(c, c->Jc + o, v1, v2, c->Ls + o);
            }
            /* DSS of the Laplacians (weighted_dss! pairs, hyperdiffusion.jl:373-406) */
            dss_vector(c, c->L1, c->L2, nl);
            dss_scalar(c, c->L3, nl);
            dss_scalar(c, c->Ls, nl);
            /* apply: hyperdiffusion.jl:273-307 */
            real nv = (real)c->n.nu4_vort, ns = (real)c->n.nu4_scal, dd = (real)c->n.div_damp;
            for (int k = 0; k < nl; ++k) {
                int64_t o = (int64_t)k * nc;
                real *gd1 = c->t[8], *gd2 = c->t[8] + nc, *wc1 = c->t[9], *wc2 = c->t[9] + nc, *wc3 = c->t[10];
                vector_laplacian_parts(c, k, c->L1 + o, c->L2 + o, c->L3 + o, gd1, gd2, wc1, wc2, wc3);
                ORA_PAR_FOR
                for (int col = 0; col < nc; ++col) {
                    real d1 = dd * gd1[col] - wc1[col];
                    real d2 = dd * gd2[col] - wc2[col];
                    real d3 = dd * 0 - wc3[col];
                    UH1(Yt)[o + col] -= nv * d1;
                    UH2(Yt)[o + col] -= nv * d2;
                    c->t[11][o + col] = d3; /* kept for the face interpolation */
                }
                real *q1 = c->t[0], *q2 = c->t[0] + nc, *v1 = c->t[1], *v2 = c->t[1] + nc, *w = c->t[2];
                op_grad(c, c->Ls + o, q1, q2);
                ORA_PAR_FOR
                for (int col = 0; col < nc; ++col) {
                    real a1, a2;
                    to_contra(c, col, c->sc[k], q1[col], q2[col], &a1, &a2);
                    v1[col] = rho[o + col] * a1;
                    v2[col] = rho[o + col] * a2;
                }
                op_wdiv(c, c->Jc + o, v1, v2, w);
                ORA_PAR_FOR
                for (int col = 0; col < nc; ++col) RHOE(Yt)[o + col] -= ns * w[col];
            }
            /* u_3 -= nu_v * winterp(J rho, C3(nabla^4 u)) */
            ORA_PAR_FOR
            for (int col = 0; col < nc; ++col)
                for (int f = 0; f <= nl; ++f) {
                    real v;
                    if (f == 0) v = c->t[11][IDX(0, col)];
                    else if (f == nl) v = c->t[11][IDX(nl - 1, col)];
                    else {
                        int64_t a = IDX(f - 1, col), b = IDX(f, col);
                        real wa = c->Jc[a] * rho[a], wb = c->Jc[b] * rho[b];
                        v = ((wa * c->t[11][a] + wb * c->t[11][b]) / 2) / ((wa + wb) / 2);
                    }
                    U3(Yt)[IDX(f, col)] -= nv * v;
                }
        }
    }
    (void)e4;
    /* ---- explicit_vertical_advection_tendency!, advection.jl:205-290 ---- */
    /* omega^3 = wcurl(u_h) (3-D only) */
    for (int k = 0; k < nl; ++k) {
        int64_t o = (int64_t)k * nc;
        if (sphere) op_wcurl3(c, c->Jc + o, u1 + o, u2 + o, om3 + o);
        else
            ORA_PAR_FOR
            for (int col = 0; col < nc; ++col) om3[o + col] = 0;
    }
    /* omega^12 at faces: curl_v(u_h) + CT12(wcurl(u_3)) */
    real *w12a = c->t[6], *w12b = c->t[7];
    for (int f = 0; f <= nl; ++f) {
        int64_t o = (int64_t)f * nc;
        ORA_PAR_FOR
        for (int col = 0; col < nc; ++col) {
            real a = 0, b = 0;
            if (f >= 1 && f <= nl - 1) {
                real Jf = c->Jf[o + col];
                a = -(u2[IDX(f, col)] - u2[IDX(f - 1, col)]) / Jf;
                b = (u1[IDX(f, col)] - u1[IDX(f - 1, col)]) / Jf;
            }
            w12a[o + col] = a;
            w12b[o + col] = b;
        }
        if (sphere) {
            real *x = c->t[4], *y = c->t[4] + nc;
            op_wcurl12(c, c->Jf + o, u3 + o, x, y);
            ORA_PAR_FOR
            for (int col = 0; col < nc; ++col) {
                w12a[o + col] += x[col];
                w12b[o + col] += y[col];
            }
        }
    }
    const double *G = c->g.gup;
    /* face C12 quantity  (f12 + omega12) x (I^f(rho J) u^3), and the C3 one */
    real *Ga = c->t[4], *Gb = c->t[5];
    real *ct1 = c->t[2], *ct2 = c->t[3]; /* CT12(c_u) at centres */
    for (int k = 0; k < nl; ++k)
        ORA_PAR_FOR
        for (int col = 0; col < nc; ++col) {
            int64_t q = IDX(k, col);
            to_contra(c, col, c->sc[k], u1[q], u2[q], &ct1[q], &ct2[q]);
        }
    (void)G;
    for (int f = 0; f <= nl; ++f)
        ORA_PAR_FOR
        for (int col = 0; col < nc; ++col) {
            int64_t q = IDX(f, col);
            real sf = c->sf[f];
            real fa = (real)c->g.f12[col] / sf, fb = (real)c->g.f12[nc + col] / sf;
            real A1 = fa + w12a[q], A2 = fb + w12b[q];
            real rJ;
            if (f == 0) rJ = rho[IDX(0, col)] * c->Jc[IDX(0, col)];
            else if (f == nl) rJ = rho[IDX(nl - 1, col)] * c->Jc[IDX(nl - 1, col)];
            else rJ = (rho[IDX(f - 1, col)] * c->Jc[IDX(f - 1, col)] + rho[IDX(f, col)] * c->Jc[IDX(f, col)]) / 2;
            real b3 = rJ * c->fu3[q];
            real Jf = c->Jf[q];
            Ga[q] = Jf * (A2 * b3);
            Gb[q] = Jf * (0 - A1 * b3);
            /* u_3 tendency: (f12 + omega12) x I^f(CT12(c_u)) + grad_v(K) */
            real i1, i2;
            if (f == 0) {
                i1 = ct1[IDX(0, col)];
                i2 = ct2[IDX(0, col)];
            } else if (f == nl) {
                i1 = ct1[IDX(nl - 1, col)];
                i2 = ct2[IDX(nl - 1, col)];
            } else {
                i1 = (ct1[IDX(f - 1, col)] + ct1[IDX(f, col)]) / 2;
                i2 = (ct2[IDX(f - 1, col)] + ct2[IDX(f, col)]) / 2;
            }
            real cr3 = Jf * (A1 * i2 - A2 * i1);
            real gK = (f >= 1 && f <= nl - 1) ? c->K[IDX(f, col)] - c->K[IDX(f - 1, col)] : 0;
            U3(Yt)[q] -= cr3 + gK;
        }
    for (int k = 0; k < nl; ++k)
        ORA_PAR_FOR
        for (int col = 0; col < nc; ++col) {
            int64_t q = IDX(k, col);
            real rJ = rho[q] * c->Jc[q];
            real ia = (Ga[IDX(k, col)] + Ga[IDX(k + 1, col)]) / 2;
            real ib = (Gb[IDX(k, col)] + Gb[IDX(k + 1, col)]) / 2;
            real a3 = (real)c->g.f3_num[col] / c->Jvc[k] + om3[q];
            real J = c->Jc[q];
            real x1 = J * (0 - a3 * ct2[q]);
            real x2 = J * (a3 * ct1[q] - 0);
            UH1(Yt)[q] -= ia / rJ + x1;
            UH2(Yt)[q] -= ib / rJ + x2;
        }
    /* ---- additional_tendency! (dry subset), remaining_tendency.jl:106-267 ---- */
    if (c->n.rayleigh && c->n.alpha_uh != 0)
        for (int k = 0; k < nl; ++k) {
            real beta = (real)c->n.alpha_uh * rayleigh_zeta(c, c->zc[k]);
            ORA_PAR_FOR
            for (int col = 0; col < nc; ++col) {
                UH1(Yt)[IDX(k, col)] += -beta * u1[IDX(k, col)];
                UH2(Yt)[IDX(k, col)] += -beta * u2[IDX(k, col)];
            }
        }
    if (c->n.vert_diff && !c->n.implicit_diffusion) vertical_diffusion_tendency(c, Yt, Y);
    if (c->n.sfc_flux) surface_flux_tendency(c, Yt, Y);
}
void ora_remaining_tendency(ora_ctx *c, ora_real *Yt, ora_real *Y) { remaining_tendency(c, (real *)Yt, (real *)Y); }

/* ------------------------------------------------------------ IMEX step */
/* ARS(3,4,3) of Ascher, Ruuth & Spiteri (1997), with the explicit a42 = a43
 * family; CTS IMEXAlgorithm(ARS343, NewtonsMethod(max_iters = 1)) drives the
 * hooks as in integrator.jl:193-225: per stage the stage state is DSSed,
 * one Newton iteration is taken with the Jacobian refreshed, the post-Newton
 * upwind correction is applied after a cache_imp! refresh, the implicit
 * tendency is recovered as (U - U~) / (dt gamma), the cache is refreshed at
 * the end of the stage, and the explicit tendency is evaluated.  The step
 * ends with u + dt sum b (F_E + F_I), a DSS and a cache refresh. */
typedef struct {
    double aE[4][4], aI[4][4], b[4], c[4];
} ars343_t;
static ars343_t ars343(void) {
    ars343_t t;
    memset(&t, 0, sizeof(t));
    /* gamma: root of 6 g^3 - 18 g^2 + 9 g - 1 = 0 in (0, 1/2) */
    double g = 0.4358665215;
    for (int it = 0; it < 50; ++it) g -= (6 * g * g * g - 18 * g * g + 9 * g - 1) / (18 * g * g - 36 * g + 9);
    double b1 = -1.5 * g * g + 4 * g - 0.25, b2 = 1.5 * g * g - 5 * g + 1.25;
    double a42 = 0.5529291480359398, a43 = a42;
    double a31 = (1 - 4.5 * g + 1.5 * g * g) * a42 + (2.75 - 10.5 * g + 3.75 * g * g) * a43 - 3.5 + 13 * g - 4.5 * g * g;
    double a32 = (-1 + 4.5 * g - 1.5 * g * g) * a42 + (-2.75 + 10.5 * g - 3.75 * g * g) * a43 + 4 - 12.5 * g + 4.5 * g * g;
    double a41 = 1 - a42 - a43;
    t.aE[1][0] = g;
    t.aE[2][0] = a31;
    t.aE[2][1] = a32;
    t.aE[3][0] = a41;
    t.aE[3][1] = a42;
    t.aE[3][2] = a43;
    t.aI[1][1] = g;
    t.aI[2][1] = (1 - g) / 2;
    t.aI[2][2] = g;
    t.aI[3][1] = b1;
    t.aI[3][2] = b2;
    t.aI[3][3] = g;
    t.b[0] = 0;
    t.b[1] = b1;
    t.b[2] = b2;
    t.b[3] = g;
    t.c[0] = 0;
    t.c[1] = g;
    t.c[2] = (1 + g) / 2;
    t.c[3] = 1;
    return t;
}

/* Storage rounding: with it on, ora_step rounds every array a float32
 * implementation stores (each stage state, explicit and implicit tendency,
 * and the new state) to float32, and computes in the build's precision
 * between the roundings.  It measures the error float32 storage alone
 * causes (the float64 build only; no effect in the float32 build). */
void ora_set_storage_rounding(ora_ctx *c, int32_t on) { c->round_storage = on; }
static void round_storage(const ora_ctx *c, real *x, int64_t n) {
    if (!c->round_storage) return;
    ORA_PAR_FOR
    for (int64_t q = 0; q < n; ++q) x[q] = (real)(float)x[q];
}

void ora_step(ora_ctx *c, ora_real *Yd, double dt, int32_t nsteps) {
    real *u = (real *)Yd;
    ars343_t tab = ars343();
    int64_t S = c->S;
    for (int st = 0; st < nsteps; ++st) {
        /* stage 1 (explicit): U1 = u */
        set_precomputed(c, u);
        remaining_tendency(c, c->FE[0], u);
        round_storage(c, c->FE[0], S);
        for (int i = 1; i < 4; ++i) {
            real *Ut = c->Ut, *U = c->U;
            ORA_PAR_FOR
            for (int64_t q = 0; q < S; ++q) {
                real s = 0;
                for (int j = 0; j < i; ++j) {
                    if (tab.aE[i][j] != 0) s += (real)tab.aE[i][j] * c->FE[j][q];
                    if (j >= 1 && tab.aI[i][j] != 0) s += (real)tab.aI[i][j] * c->FI[j][q];
                }
                Ut[q] = u[q] + (real)dt * s;
            }
            dss_state(c, Ut);
            round_storage(c, Ut, S);
            real dtg = (real)(dt * tab.aI[i][i]);
            memcpy(U, Ut, sizeof(real) * S);
            /* Newton iteration: cache_imp!, T_imp!, Wfact, ldiv! */
            set_precomputed(c, U);
            implicit_tendency(c, c->Rv, U);
            ORA_PAR_FOR
            for (int64_t q = 0; q < S; ++q) c->Rv[q] = Ut[q] + dtg * c->Rv[q] - U[q];
            update_jacobian(c, U, dtg);
            if (c->n.vert_diff && c->n.implicit_diffusion)
                jac_ldiv_approx(c, c->dY, c->Rv);
            else
                jac_ldiv(c, c->dY, c->Rv);
            ORA_PAR_FOR
            for (int64_t q = 0; q < S; ++q) U[q] -= c->dY[q];
            /* post-Newton correction (T_post_imp!) after a cache_imp! refresh */
            if (c->n.upwinding != 0) {
                set_precomputed(c, U);
                correct_implicit(c, c->dY, U);
                ORA_PAR_FOR
                for (int64_t q = 0; q < S; ++q) U[q] += dtg * c->dY[q];
            }
            round_storage(c, U, S);
            ORA_PAR_FOR
            for (int64_t q = 0; q < S; ++q) c->FI[i][q] = (U[q] - Ut[q]) / dtg;
            round_storage(c, c->FI[i], S);
            /* end of stage: cache! then T_exp! */
            set_precomputed(c, U);
            remaining_tendency(c, c->FE[i], U);
            round_storage(c, c->FE[i], S);
        }
        ORA_PAR_FOR
        for (int64_t q = 0; q < S; ++q) {
            real s = 0;
            for (int j = 0; j < 4; ++j)
                if (tab.b[j] != 0) s += (real)tab.b[j] * (c->FE[j][q] + (j >= 1 ? c->FI[j][q] : 0));
            u[q] = u[q] + (real)dt * s;
        }
        dss_state(c, u);
        round_storage(c, u, S);
        set_precomputed(c, u);
    }
}

/* ------------------------------------------------- operator entry points */
#define OPWRAP_BEGIN                                                   \
    int nc = c->ncol;                                                  \
    real *a = (real *)calloc((size_t)nc * 6, sizeof(real));                   \
    real *b = a + nc, *x = a + 2 * nc, *y = a + 3 * nc, *z = a + 4 * nc; \
    (void)b; (void)x; (void)y; (void)z;
#define OPWRAP_END free(a);

void ora_op_grad(ora_ctx *c, int k, const double *f, double *o1, double *o2) {
    OPWRAP_BEGIN
    (void)k;
    for (int i = 0; i < nc; ++i) a[i] = (real)f[i];
    op_grad(c, a, x, y);
    for (int i = 0; i < nc; ++i) { o1[i] = x[i]; o2[i] = y[i]; }
    OPWRAP_END
}
void ora_op_wgrad(ora_ctx *c, int k, const double *f, double *o1, double *o2) {
    OPWRAP_BEGIN
    (void)k;
    for (int i = 0; i < nc; ++i) a[i] = (real)f[i];
    op_wgrad(c, a, x, y);
    for (int i = 0; i < nc; ++i) { o1[i] = x[i]; o2[i] = y[i]; }
    OPWRAP_END
}
void ora_op_div(ora_ctx *c, int k, const double *v1, const double *v2, double *o) {
    OPWRAP_BEGIN
    for (int i = 0; i < nc; ++i) { a[i] = (real)v1[i]; b[i] = (real)v2[i]; }
    op_div(c, c->Jc + (int64_t)k * nc, a, b, x);
    for (int i = 0; i < nc; ++i) o[i] = x[i];
    OPWRAP_END
}
void ora_op_wdiv(ora_ctx *c, int k, const double *v1, const double *v2, double *o) {
    OPWRAP_BEGIN
    for (int i = 0; i < nc; ++i) { a[i] = (real)v1[i]; b[i] = (real)v2[i]; }
    op_wdiv(c, c->Jc + (int64_t)k * nc, a, b, x);
    for (int i = 0; i < nc; ++i) o[i] = x[i];
    OPWRAP_END
}
void ora_op_curl3(ora_ctx *c, int k, const double *u1, const double *u2, double *o) {
    OPWRAP_BEGIN
    for (int i = 0; i < nc; ++i) { a[i] = (real)u1[i]; b[i] = (real)u2[i]; }
    op_curl3(c, c->Jc + (int64_t)k * nc, a, b, x);
    for (int i = 0; i < nc; ++i) o[i] = x[i];
    OPWRAP_END
}
void ora_op_wcurl3(ora_ctx *c, int k, const double *u1, const double *u2, double *o) {
    OPWRAP_BEGIN
    for (int i = 0; i < nc; ++i) { a[i] = (real)u1[i]; b[i] = (real)u2[i]; }
    op_wcurl3(c, c->Jc + (int64_t)k * nc, a, b, x);
    for (int i = 0; i < nc; ++i) o[i] = x[i];
    OPWRAP_END
}
void ora_op_curl12(ora_ctx *c, int k, const double *u3, double *o1, double *o2) {
    OPWRAP_BEGIN
    for (int i = 0; i < nc; ++i) a[i] = (real)u3[i];
    op_curl12(c, c->Jc + (int64_t)k * nc, a, x, y);
    for (int i = 0; i < nc; ++i) { o1[i] = x[i]; o2[i] = y[i]; }
    OPWRAP_END
}
void ora_op_wcurl12(ora_ctx *c, int k, const double *u3, double *o1, double *o2) {
    OPWRAP_BEGIN
    for (int i = 0; i < nc; ++i) a[i] = (real)u3[i];
    op_wcurl12(c, c->Jc + (int64_t)k * nc, a, x, y);
    for (int i = 0; i < nc; ++i) { o1[i] = x[i]; o2[i] = y[i]; }
    OPWRAP_END
}

void ora_tableau(double *aE, double *aI, double *b, double *cc) {
    ars343_t t = ars343();
    for (int i = 0; i < 4; ++i) {
        for (int j = 0; j < 4; ++j) {
            aE[4 * i + j] = t.aE[i][j];
            aI[4 * i + j] = t.aI[i][j];
        }
        b[i] = t.b[i];
        cc[i] = t.c[i];
    }
}
