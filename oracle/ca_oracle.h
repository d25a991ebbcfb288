/*
 * ca_oracle.h -- CPU restatement of one ClimaAtmos dry-dycore IMEX step.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker the native HIP library is
 * compared against (and, built in float with OpenMP, the CPU timing baseline).
 * Nothing in the product links or calls it.
 *
 * Every function restates the reference's algorithm, operator by operator,
 * in the order the reference materialises its broadcasts; the file:line it
 * follows is cited at each function in ca_oracle.c.
 */
#ifndef CA_ORACLE_H
#define CA_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ora_geom {
    int32_t nelem;        /* horizontal elements; 0 => independent columns   */
    int32_t ncol;         /* stored columns (16 * nelem, or column count)    */
    int32_t nlev;         /* centre levels                                   */
    double radius;        /* deep-atmosphere radius, 0 => no s(z) scaling    */
    const double *z_face; /* nlev + 1                                        */
    const double *J_h;    /* ncol                                            */
    const double *gup;    /* 3 * ncol: g^11, g^12, g^22 (plane-major)        */
    const double *gdn;    /* 3 * ncol: g_11, g_12, g_22                      */
    const double *f3_num; /* ncol                                            */
    const double *f12;    /* 2 * ncol                                        */
    const double *basis;  /* 12 * ncol: a^1, a^2, a_1, a_2 (x,y,z each)      */
    int32_t n_dss;        /* shared-node groups                              */
    const int32_t *dss_ptr;
    const int32_t *dss_col;
    const double *dss_w;
    const double *gll_w;  /* 4 */
    const double *gll_D;  /* 16, D[i*4+k] = l_k'(xi_i) */
} ora_geom;

typedef struct ora_num {
    double R_d, cp_d, cv_d, T_0, p_ref_theta, grav, T_surf_ref, T_min_ref, T_min_sgs;
    int32_t hyperdiff;
    double nu4_vort, nu4_scal, div_damp;
    int32_t upwinding;     /* 0 none, 1 first_order, 3 vanleer_limiter */
    int32_t rayleigh;
    double zd_rayleigh, alpha_uh, alpha_w;   /* defaults for all columns     */
    const double *alpha_w_col;               /* optional per-column alpha_w   */
    int32_t vert_diff;
    const double *D0_col;                    /* per-column diffusivity [m2/s] */
    double dt;             /* p.dt, used by the van Leer limiter */
    int32_t sfc_flux;      /* surface_flux_tendency! (ExchangeCoefficients)  */
    double Cd, Ch;         /* drag and heat exchange coefficients            */
    double gustiness;      /* wind-speed floor [m/s]                         */
    const double *T_sfc_col; /* per-column surface temperature [K]         */
    int32_t implicit_diffusion; /* vertical diffusion in T_imp (diff_mode Implicit) */
    int32_t solve_iters;   /* approximate_solve_iters of the linear solve    */
    int32_t vdiff_kind;    /* 0 DecayWithHeightDiffusion, 1 VerticalDiffusion */
    double vdiff_H;        /* decay height [m] (inf: K = D0)                 */
    double vdiff_C_E;      /* VerticalDiffusion's bulk coefficient           */
} ora_num;

#ifdef ORACLE_F32
typedef float ora_real;
#else
typedef double ora_real;
#endif

typedef struct ora_ctx ora_ctx;

int ora_create(const ora_geom *g, const ora_num *n, ora_ctx **out);
void ora_destroy(ora_ctx *c);
int64_t ora_state_size(const ora_ctx *c);

/* hooks (state vectors are flat, native layout [rho|u1|u2|rhoe|u3]) */
void ora_set_precomputed(ora_ctx *c, ora_real *Y);            /* filters Y */
void ora_remaining_tendency(ora_ctx *c, ora_real *Yt, ora_real *Y);
void ora_implicit_tendency(ora_ctx *c, ora_real *Yt, ora_real *Y);
void ora_update_jacobian(ora_ctx *c, ora_real *Y, double dtgamma);
void ora_ldiv(ora_ctx *c, ora_real *dY, const ora_real *R);
void ora_correct_implicit(ora_ctx *c, ora_real *Yt, ora_real *Y);
void ora_dss(ora_ctx *c, ora_real *Y);
void ora_step(ora_ctx *c, ora_real *Y, double dt, int32_t nsteps);
/* round every array a float32 implementation stores to float32 inside
 * ora_step (stage states, tendencies, new state): measures the error of
 * float32 storage alone */
void ora_set_storage_rounding(ora_ctx *c, int32_t on);
/* the ARS343 tableau used by ora_step (4x4 row-major aE, aI; b; c) */
void ora_tableau(double *aE, double *aI, double *b, double *c);

/* debug access: precomputed centre fields (K, T, p, h_tot) and the last
 * hyperdiffusion Laplacians (4 centre planes: L1, L2, L3, Ls) */
void ora_get_precomputed(const ora_ctx *c, double *K, double *T, double *p, double *h);
void ora_get_laplacians(const ora_ctx *c, double *out);

/* operator-level entry points used by the pinning tests (one level, all
 * elements; inputs/outputs are ncol arrays) */
void ora_op_grad(ora_ctx *c, int k, const double *f, double *o1, double *o2);
void ora_op_wgrad(ora_ctx *c, int k, const double *f, double *o1, double *o2);
void ora_op_div(ora_ctx *c, int k, const double *v1, const double *v2, double *o);
void ora_op_wdiv(ora_ctx *c, int k, const double *v1, const double *v2, double *o);
void ora_op_curl3(ora_ctx *c, int k, const double *u1, const double *u2, double *o);
void ora_op_wcurl3(ora_ctx *c, int k, const double *u1, const double *u2, double *o);
void ora_op_curl12(ora_ctx *c, int k, const double *u3, double *o1, double *o2);
void ora_op_wcurl12(ora_ctx *c, int k, const double *u3, double *o1, double *o2);
/* dense column Jacobian of implicit_tendency by central differences, and the
 * band blocks of update_jacobian, for one column (pinning tests) */
void ora_column_blocks(ora_ctx *c, int col, double *Aru, double *Aeu, double *Aur,
                       double *Aue, double *Auu_lo, double *Auu_d, double *Auu_up);
/* the implicit-diffusion blocks of one column: (rho e_tot, rho e_tot) and
 * (u_h, u_h), tridiagonal on centres (lo, d, up each nlev) */
void ora_column_diffusion_blocks(ora_ctx *c, int col, double *Aee, double *Ahh);
/* the (u_3, u_h) blocks of one column: faces x (centre f-1: u_1, u_2; centre f: u_1, u_2) */
void ora_column_uh_blocks(ora_ctx *c, int col, double *Auh);

#ifdef __cplusplus
}
#endif
#endif
