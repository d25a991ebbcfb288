/*
 * cadycore.h -- C ABI of the MI355X-native ClimaAtmos dry-dycore step.
 *
 * The entry points replace, one for one, the hooks ClimaTimeSteppers calls on
 * the ClimaAtmos ODE function built in src/simulation/integrator.jl:193-225:
 *
 *   ca_remaining_tendency          remaining_tendency!(Yt, Yt_lim, Y, p, t)
 *                                  src/prognostic_equations/remaining_tendency.jl:48
 *   ca_implicit_tendency           implicit_tendency!(Yt, Y, p, t)
 *                                  src/prognostic_equations/implicit/implicit_tendency.jl:36
 *   ca_update_jacobian             update_jacobian!(jac, Y, p, dtgamma, t)   (Wfact)
 *                                  src/prognostic_equations/implicit/jacobian.jl:74-75
 *   ca_ldiv                        LinearAlgebra.ldiv!(dY, jac, R)
 *                                  src/prognostic_equations/implicit/jacobian.jl:78-82
 *   ca_correct_implicit_advection  correct_implicit_advection_tendency!(Yt, Y, p, t) (T_post_imp!)
 *                                  src/prognostic_equations/implicit/implicit_tendency.jl:322
 *   ca_set_precomputed             set_implicit_precomputed_quantities!(Y, p, t)  (cache_imp!)
 *                                  src/cache/precomputed_quantities.jl:698
 *   ca_dss                         dss!(Y, p, t)  src/prognostic_equations/constrain_state.jl:59
 *   ca_step                        one IMEXAlgorithm(ARS343, NewtonsMethod(max_iters = 1)) step,
 *                                  as driven by CTS.step! from solve_atmos!
 *                                  (src/simulation/solve.jl:119-158)
 *
 * All state pointers are DEVICE pointers to float32 data in the native
 * layout: [rho | u_1 | u_2 | rho e_tot] centre planes (nlev x ncol each,
 * level-major) followed by the u_3 face plane ((nlev + 1) x ncol), i.e. the
 * reference's Y.c = (rho, u_h::C12, rho e_tot) and Y.f = (u_3::C3).
 * ca_pack_state / ca_unpack_state convert from / to any strided layout
 * (ClimaCore's VIJFH / VIJHF parent arrays included).
 *
 * Calls are stream-ordered on the hipStream_t passed as `stream` (NULL = the
 * default stream).  A handle is not thread-safe: the caller serialises calls
 * on one handle, as ClimaTimeSteppers does.  Nothing throws across the ABI;
 * every call returns a ca_status and ca_last_error() describes the last
 * failure.
 */
#ifndef CADYCORE_H
#define CADYCORE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum ca_status {
    CA_OK = 0,
    CA_ERR_ARG = 1,        /* invalid argument (sizes, null pointers, unsupported option) */
    CA_ERR_DEVICE = 2,     /* a HIP runtime call failed */
    CA_ERR_NONFINITE = 3,  /* a state check found NaN or Inf */
} ca_status;

typedef struct ca_handle ca_handle;

/* Geometry, factored into a horizontal part per stored column and a vertical
 * part per level (no topography).  Column index = 16 * element + 4 * j + i.
 * A Julia caller fills it from ClimaCore's LocalGeometry at one level and the
 * face heights; cadycore/geometry.py builds it for a standalone run. */
typedef struct ca_geometry {
    int32_t nelem;          /* horizontal elements (0 => independent columns) */
    int32_t ncol;           /* stored columns                                */
    int32_t nlev;           /* centre levels                                 */
    double radius;          /* deep-atmosphere radius (0 => no scaling)      */
    const double *z_face;   /* [nlev + 1] face heights                       */
    const double *J_h;      /* [ncol] horizontal Jacobian at radius          */
    const double *gup;      /* [3 * ncol] g^11, g^12, g^22 at radius         */
    const double *gdn;      /* [3 * ncol] g_11, g_12, g_22 at radius         */
    const double *f3_num;   /* [ncol] 2 Omega sin(lat) (f-plane: f)          */
    const double *f12;      /* [2 * ncol] contravariant horizontal Coriolis  */
    const double *basis;    /* [12 * ncol] Cartesian a^1, a^2, a_1, a_2      */
    int32_t n_dss;          /* number of shared-node groups                  */
    const int32_t *dss_ptr; /* [n_dss + 1]                                   */
    const int32_t *dss_col; /* [dss_ptr[n_dss]] member columns               */
    const double *dss_w;    /* [dss_ptr[n_dss]] weights J / sum J            */
    const double *gll_w;    /* [4] GLL weights                               */
    const double *gll_D;    /* [16] D[i*4+k] = l_k'(xi_i)                    */
} ca_geometry;

/* Thermodynamic constants (ClimaParams names in cadycore/ca_params.toml).
 * Dry air: R_d / cp_d (kappa_d) must be 2/7 to within 1e-6, else ca_create
 * and ca_ensemble_create return CA_ERR_ARG. */
typedef struct ca_params {
    double R_d, cp_d, cv_d, T_0, p_ref_theta, grav, T_surf_ref, T_min_ref, T_min_sgs;
} ca_params;

/* Numerics of the dry parity profile. */
typedef struct ca_config {
    int32_t hyperdiff;      /* 1: Hyperdiffusion (sphere only)               */
    double nu4_vort, nu4_scal, div_damp;
    int32_t upwinding;      /* 0 none, 1 first_order, 3 vanleer_limiter      */
    int32_t rayleigh;       /* 1: RayleighSponge                             */
    double zd_rayleigh, alpha_uh, alpha_w;
    const double *alpha_w_col; /* optional [ncol] per-column alpha_w        */
    int32_t vert_diff;      /* 1: DecayWithHeightDiffusion(D0 per column, H = Inf) */
    const double *D0_col;   /* [ncol] when vert_diff                         */
    double dt;              /* p.dt (van Leer Courant number)                */
    /* surface_flux_tendency! (src/prognostic_equations/surface_flux.jl:101-145)
     * with the default AtmosSurface (src/types.jl:2292-2302): bulk fluxes of
     * ExchangeCoefficients(Cd, Ch) against a per-column surface temperature
     * (the AnalyticTemperature evaluated at the surface), dry air, no
     * boundary overrides.  Sphere and independent columns; the reverse mode
     * and the column ensembles return CA_ERR_ARG when it is on. */
    int32_t sfc_flux;       /* 1: surface fluxes on (0: disable_surface_flux_tendency) */
    double Cd, Ch;          /* exchange coefficients (default 0.0044 each)   */
    double gustiness;       /* wind-speed floor [m/s] (default 1)            */
    const double *T_sfc_col; /* [ncol] surface temperature [K] when sfc_flux */
    /* diff_mode Implicit (YAML implicit_diffusion: true): the vertical
     * diffusion moves to the implicit tendency, the Jacobian gains its
     * (rho e_tot, rho e_tot) and (u_h, u_h) blocks (update_diffusion_jacobian!,
     * manual_sparse_jacobian.jl:1031-1261) and ldiv becomes the approximate
     * block arrowhead solve with solve_iters iterations (:538-578).  Needs
     * vert_diff; the reverse mode and the column ensembles return CA_ERR_ARG. */
    int32_t implicit_diffusion;
    int32_t solve_iters;    /* approximate_linear_solve_iters, 1 .. 100 (default 1) */
    /* The vertical-diffusion closure (vert_diff, src/config/model_getters.jl:341-355;
     * ᶜcompute_eddy_diffusivity_coefficient, cache/eddy_diffusivity_coefficient.jl:16-41):
     *   0 DecayWithHeightDiffusion: K = D0_col exp(-(z - z_sfc) / vert_diff_H)
     *     (vert_diff_H = 0 or INFINITY: K = D0_col, the zero-initialised default);
     *   1 VerticalDiffusion: K = vert_diff_C_E |u_h| dz_1 / 2, |u_h| the lowest
     *     centre's wind speed and dz_1 the lowest layer's depth, times
     *     exp(-((85000 - p) / 10000)^2) where p < 85000 Pa (no D0_col).
     * The face diffusivity is the harmonic mean of the centre values.  The
     * reverse mode and the column ensembles take closure 0 with H = INFINITY
     * only (CA_ERR_ARG otherwise). */
    int32_t vert_diff_model;
    double vert_diff_H;     /* decay height [m] (0 or INFINITY: no decay)     */
    double vert_diff_C_E;   /* bulk coefficient of VerticalDiffusion, >= 0    */
} ca_config;

/* Strided layout of one component: element (level v, node i, j, element h)
 * lives at base[offset + v*s_v + i*s_i + j*s_j + h*s_h] (float units). */
typedef struct ca_layout {
    int64_t offset, s_v, s_i, s_j, s_h;
} ca_layout;

/* ca_create checks its arguments and returns CA_ERR_ARG (message in
 * ca_last_error(NULL)) on: null required tables (z_face, J_h, gup, gdn, f3_num,
 * f12, gll_w, gll_D; on a sphere also basis and the dss_* arrays), nlev < 3,
 * ncol outside [1, 2^28], ncol != 16 nelem on a sphere, z_face not strictly
 * increasing, GLL tables other than the Nq = 4 Gauss-Lobatto-Legendre ones
 * (the kernels share them process-wide), an unknown upwinding code, vertical
 * diffusion without D0_col, surface fluxes without T_sfc_col or with negative
 * Cd, Ch or gustiness, or R_d / cp_d off 2/7. */
ca_status ca_create(const ca_geometry *geom, const ca_params *params, const ca_config *config,
                    ca_handle **out);
ca_status ca_destroy(ca_handle *h);
int64_t ca_state_size(const ca_handle *h);
/* message of the last failure on h; with h = NULL, of the last failed
 * ca_create (which frees everything it allocated and leaves *out NULL) */
const char *ca_last_error(const ca_handle *h);

/* hook-level entry points (device pointers, stream-ordered) */
ca_status ca_set_precomputed(ca_handle *h, float *Y, void *stream);
ca_status ca_remaining_tendency(ca_handle *h, float *Yt, const float *Y, void *stream);
ca_status ca_implicit_tendency(ca_handle *h, float *Yt, const float *Y, void *stream);
ca_status ca_update_jacobian(ca_handle *h, const float *Y, double dtgamma, void *stream);
ca_status ca_ldiv(ca_handle *h, float *dY, const float *R, void *stream);
ca_status ca_correct_implicit_advection(ca_handle *h, float *Yt, const float *Y, void *stream);
ca_status ca_dss(ca_handle *h, float *Y, void *stream);

/* fused ARS343 step(s): Y <- step^nsteps(Y) */
ca_status ca_step(ca_handle *h, float *Y, double dt, int32_t nsteps, void *stream);
/* ca_step as a hipGraph replay.  The first call captures one step's launches
 * on `stream` into a graph the handle keeps; later calls with the same Y,
 * dt and stream replay it nsteps times, and any other Y, dt or stream (or a
 * ca_set_tuning in between) captures again.  Bitwise the results of ca_step.
 * `stream` must be a non-default stream; profiling and debug poisoning must
 * be off (CA_ERR_ARG otherwise).  On small grids, where a step is many short
 * launches, the replay removes their per-launch host cost.  No counterpart
 * in the reference (CTS.step! launches through Julia). */
ca_status ca_step_graph(ca_handle *h, float *Y, double dt, int32_t nsteps, void *stream);

/* layout conversion: 5 components (rho, u_1, u_2, rho e_tot on nlev levels;
 * u_3 on nlev + 1 levels) between strided caller arrays and native Y */
ca_status ca_pack_state(ca_handle *h, const float *src, const ca_layout *layouts5, float *Y,
                        void *stream);
ca_status ca_unpack_state(ca_handle *h, const float *Y, const ca_layout *layouts5, float *dst,
                          void *stream);

/* ------------------------------------------------------------------------
 * The hooks on ClimaCore FieldVectors.  A ca_fieldvector is a state (or a
 * tendency, increment or residual) as ClimaCore holds it: the parent array of
 * the centre fields Y.c = (rho, u_h::C12, rho e_tot) and of the face field
 * Y.f = (u_3::C3), with one ca_layout per component (rho, u_1, u_2,
 * rho e_tot relative to c; u_3 relative to f) giving its offset and its
 * level / node i / node j / element strides, so a caller passes
 * pointer(parent(Y.c)), pointer(parent(Y.f)) and the strides of its
 * DataLayout (VIJFH, VIJHF, or any other order).  Arguments, order and
 * meaning are those of the reference hooks, t included (the dry path does not
 * depend on it).  A FieldVector in the native layout (ca_fieldvector_native)
 * is read and written in place; any other layout is transposed through LDS
 * into a native staging buffer of the handle (two state-sized buffers,
 * allocated by the first such call and kept until ca_destroy) and outputs are
 * transposed back: one extra read and write of each argument.  Results are
 * bitwise those of the native-pointer hooks above. */
typedef struct ca_fieldvector {
    float *c;               /* parent array of Y.c                            */
    float *f;               /* parent array of Y.f                            */
    ca_layout lay[5];       /* rho, u_1, u_2, rho e_tot (in c); u_3 (in f)    */
} ca_fieldvector;
/* descriptor of a native state buffer Y (ca_state_size(h) floats) */
ca_status ca_fieldvector_native(const ca_handle *h, float *Y, ca_fieldvector *out);
/* descriptor of ClimaCore VIJFH parent arrays: Y.c as Julia (Nv, 4, 4, 4, Nh)
 * and Y.f as (Nv + 1, 4, 4, 1, Nh), column-major (level fastest) */
ca_status ca_fieldvector_vijfh(const ca_handle *h, float *c, float *f, ca_fieldvector *out);
/* cache!: set_precomputed_quantities!(Y, p, t), src/cache/precomputed_quantities.jl:928-931.
 * The precomputed quantities are formed inside the kernels that use them; the
 * effect a caller sees is the impenetrability filter on Y's u_3. */
ca_status ca_set_precomputed_quantities_fv(ca_handle *h, const ca_fieldvector *Y, double t, void *stream);
/* cache_imp!: set_implicit_precomputed_quantities!(Y, p, t), precomputed_quantities.jl:698 */
ca_status ca_set_implicit_precomputed_quantities_fv(ca_handle *h, const ca_fieldvector *Y, double t,
                                                    void *stream);
/* T_exp_T_lim!: remaining_tendency!(Yt, Yt_lim, Y, p, t), remaining_tendency.jl:48.
 * Yt_lim (may be NULL) is zeroed, as the reference does: the dry model has no
 * limited tracer tendencies. */
ca_status ca_remaining_tendency_fv(ca_handle *h, const ca_fieldvector *Yt, const ca_fieldvector *Yt_lim,
                                   const ca_fieldvector *Y, double t, void *stream);
/* T_imp!: implicit_tendency!(Yt, Y, p, t), implicit/implicit_tendency.jl:36 */
ca_status ca_implicit_tendency_fv(ca_handle *h, const ca_fieldvector *Yt, const ca_fieldvector *Y, double t,
                                  void *stream);
/* Wfact: update_jacobian!(jac, Y, p, dtgamma, t), implicit/jacobian.jl:74-75 */
ca_status ca_update_jacobian_fv(ca_handle *h, const ca_fieldvector *Y, double dtgamma, double t, void *stream);
/* ldiv!(dY, jac, R), implicit/jacobian.jl:78-82 (the AbstractVector method at
 * :87-96 is the same call on the FieldVector the vector wraps) */
ca_status ca_ldiv_fv(ca_handle *h, const ca_fieldvector *dY, const ca_fieldvector *R, void *stream);
/* T_post_imp!: correct_implicit_advection_tendency!(Yt, Y, p, t), implicit_tendency.jl:322 */
ca_status ca_correct_implicit_advection_fv(ca_handle *h, const ca_fieldvector *Yt, const ca_fieldvector *Y,
                                           double t, void *stream);
/* dss!(Y, p, t), constrain_state.jl:59 */
ca_status ca_dss_fv(ca_handle *h, const ca_fieldvector *Y, double t, void *stream);

/* Launch-schedule knobs (the autotuner's search space).  Results do not
 * depend on them beyond float32 rounding order.
 *   CA_TUNE_NEWTON:         0 = one thread per column (Thomas sweep),
 *                           1 = one wavefront per column, 16-column tiles (PCR),
 *                           2 = one wavefront per column, 32-column tiles (PCR);
 *                           1 and 2 need nlev + 1 <= 64.
 *   CA_TUNE_EXPLICIT_CHUNK: levels per work item of the explicit-tendency
 *                           kernel (0 = whole column).  Setting it sets the
 *                           chunk of ca_step's last explicit launch (which
 *                           also forms the final combination) too; by
 *                           default that launch has a chunk of its own, and
 *                           ca_get_tuning reports the other launches'.
 *   CA_TUNE_DSS_CHUNK:      levels per work item of the DSS kernel (>= 1).
 *   CA_TUNE_DSS_KIND:       0 = staged through LDS per element cluster,
 *                           1 = direct gather from global memory.  Both sum
 *                           the same copies in the same order: results are
 *                           bitwise identical.
 *   CA_TUNE_ADJ_CHUNK:      levels per work item of the explicit-tendency
 *                           reverse (>= 1; the pullbacks' D0 partial sums
 *                           follow it).
 *   CA_TUNE_LAP_CHUNK:      levels per work item of the Laplacian kernel
 *                           (>= 1).
 * The chunk defaults follow the grid size (enough work items to fill the
 * GPU on small grids). */
typedef enum ca_tune_knob {
    CA_TUNE_NEWTON = 0,
    CA_TUNE_EXPLICIT_CHUNK = 1,
    CA_TUNE_DSS_CHUNK = 2,
    CA_TUNE_DSS_KIND = 3,
    CA_TUNE_ADJ_CHUNK = 4,
    CA_TUNE_LAP_CHUNK = 5
} ca_tune_knob;
ca_status ca_set_tuning(ca_handle *h, int32_t knob, int32_t value);
ca_status ca_get_tuning(const ca_handle *h, int32_t knob, int32_t *value);

/* non-finite check (NaN poisoning counterpart): *flag = 1 if Y holds NaN/Inf.
 * Synchronises the stream. */
ca_status ca_check_finite(ca_handle *h, const float *Y, int32_t *flag, void *stream);

/* Debug poisoning (the counterpart of fill_with_nans!, utils/debug_utils.jl:67-70,
 * which fills the cache with NaN before each tendency evaluation): while on,
 * every entry point first fills the handle's work buffers that no call
 * carries over with NaN, so a value read without being written in the same
 * call shows up as NaN in the result.  The Jacobian bands are poisoned by
 * ca_update_jacobian (which rewrites them; ca_ldiv reads them), the
 * checkpoints by ca_step_record.  Results are unchanged when the library is
 * correct; off by default. */
ca_status ca_set_debug_poison(ca_handle *h, int32_t on);

/* Per-kernel timing with HIP events recorded on the launch stream around
 * every kernel launch (off by default; enabling adds event overhead).
 * ca_profile_get returns, for kernel class `id` in [0, ca_profile_count()),
 * its name, the number of launches and the summed device time since the last
 * ca_profile_enable(h, 1).  Synchronises the device. */
ca_status ca_profile_enable(ca_handle *h, int32_t enable);
int32_t ca_profile_count(void);
ca_status ca_profile_get(ca_handle *h, int32_t id, const char **name, int64_t *launches,
                         double *total_ms);

/* ------------------------------------------------------------------------
 * Column ensembles.  No reference counterpart: the reference runs one
 * ColumnGrid per simulation (src/simulation/grids.jl:110-129) and ensembles
 * as separate processes (calibration/.../run_calibration.jl:89-110).  Each
 * member is an independent column with its own levels and parameters,
 * stepped with the same ARS343 step as ca_step on a column grid.  The
 * ensemble state is ragged: member m's [rho | u_1 | u_2 | rho e_tot]
 * (nlev each) and u_3 (nlev + 1) start at ca_ensemble_member_offset(e, m).
 * The config's hyperdiffusion fields are ignored (columns have none);
 * rayleigh and vert_diff use the members' alpha_w and D0. */
typedef struct ca_ensemble ca_ensemble;
typedef struct ca_member {
    int32_t nlev;          /* centre levels, 3 .. 128                         */
    const double *z_face;  /* [nlev + 1] face heights, increasing             */
    double alpha_w;        /* Rayleigh coefficient on u_3 (1/s)               */
    double D0;             /* vertical diffusivity (m^2/s)                    */
    double f3;             /* Coriolis parameter of the column (1/s)          */
} ca_member;
ca_status ca_ensemble_create(const ca_params *params, const ca_config *config, int32_t nmembers,
                             const ca_member *members, ca_ensemble **out);
ca_status ca_ensemble_destroy(ca_ensemble *e);
int64_t ca_ensemble_state_size(const ca_ensemble *e);
int64_t ca_ensemble_member_offset(const ca_ensemble *e, int32_t member);
const char *ca_ensemble_last_error(const ca_ensemble *e);
/* nsteps steps of every member in one persistent launch */
ca_status ca_ensemble_step(ca_ensemble *e, float *Y, double dt, int32_t nsteps, void *stream);
/* the same launch captured into a hipGraph (kept and replayed while Y, dt,
 * nsteps, the schedule and the stream stay the same); stream must not be
 * NULL (the legacy stream cannot be captured) */
ca_status ca_ensemble_step_graph(ca_ensemble *e, float *Y, double dt, int32_t nsteps, void *stream);
/* schedule knobs: CA_ENS_LANES = lanes per member (64, 128, 256);
 * CA_ENS_ORDER = member-to-workgroup order (0 as given, 1 longest first) */
typedef enum ca_ens_knob { CA_ENS_LANES = 0, CA_ENS_ORDER = 1 } ca_ens_knob;
ca_status ca_ensemble_set_tuning(ca_ensemble *e, int32_t knob, int32_t value);
ca_status ca_ensemble_get_tuning(const ca_ensemble *e, int32_t knob, int32_t *value);
/* times every schedule on a device copy of Y (Y is not modified), keeps the
 * fastest; times_ms (optional, 6 entries) receives lanes 64/128/256 x order
 * 0/1 in that order */
ca_status ca_ensemble_autotune(ca_ensemble *e, const float *Y, double dt, int32_t nsteps, void *stream,
                               double *times_ms);
/* new per-member alpha_w and / or D0 (host arrays of nmembers; NULL keeps) */
ca_status ca_ensemble_set_params(ca_ensemble *e, const double *alpha_w, const double *D0);
/* Reverse mode of nsteps ensemble steps (the vector-Jacobian product of the
 * solution operator), with step checkpointing every `every` steps.  No
 * reference counterpart: the reference differentiates only its Jacobians, in
 * forward mode (implicit/autodiff_utils.jl), and calibrates without gradients.
 * Given the initial states Y0 and the cotangent YbN of the states after
 * nsteps steps, writes the cotangent of Y0 to Yb0 (ragged, like Y0) and of
 * each member's D0 and alpha_w to D0b, awb (device arrays of nmembers).
 * Recomputes the forward trajectory itself; runs with 64 lanes per member
 * when every member has at most 32 levels or CA_ENS_LANES is 64, else 128. */
ca_status ca_ensemble_adjoint(ca_ensemble *e, const float *Y0, const float *YbN, double dt, int32_t nsteps,
                              int32_t every, float *Yb0, float *D0b, float *awb, void *stream);
/* ca_ensemble_adjoint without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * hipGraph capture, tight serving loops, harnesses that forbid hipMalloc).
 * The checkpoints and segment states live in a file-scope device array of
 * 2^24 floats (64 MiB) per GPU, held from the library's first use on that GPU
 * until the process ends; it needs (ceil(nsteps / every) + 1 + every) *
 * ca_ensemble_state_size(e) floats, and a larger request returns CA_ERR_ARG
 * before any launch.  Every call on one GPU shares that array: calls that
 * overlap on one GPU (on different streams, or from different ensembles or
 * threads) race.  Same kernels and results as ca_ensemble_adjoint. */
ca_status ca_ensemble_adjoint_noalloc(ca_ensemble *e, const float *Y0, const float *YbN, double dt, int32_t nsteps,
                                      int32_t every, float *Yb0, float *D0b, float *awb, void *stream);

/* ------------------------------------------------------------------------
 * Reverse mode of the sphere step (sphere grids; independent columns use
 * ca_ensemble_adjoint).  No reference counterpart: the reference has no
 * adjoint (forward-mode AD for its Jacobians only, implicit/autodiff_utils.jl;
 * gradient-free calibration in calibration/).  These are the pullbacks a
 * ChainRulesCore rrule or a torch.autograd.Function binds.  Device pointers;
 * the cotangents have the state's layout.
 * ------------------------------------------------------------------------ */
/* Yb <- DSS^T Yb, in place: the transpose of ca_dss (dss!, constrain_state.jl:59) */
ca_status ca_dss_vjp(ca_handle *h, float *Yb, void *stream);
/* Yb <- (d T_exp / d Y)^T Ytb at Y: the pullback of ca_remaining_tendency
 * (remaining_tendency!, remaining_tendency.jl:48), Laplacian DSS included.
 * With vertical diffusion and D0b != NULL, D0b [ncol] <- d <Ytb, T_exp> / d D0. */
ca_status ca_remaining_tendency_vjp(ca_handle *h, float *Yb, const float *Ytb, const float *Y, float *D0b,
                                    void *stream);
/* The sphere reverse mode (ca_step_record, ca_step_pullback, ca_step_adjoint)
 * supports nlev <= 128, no surface fluxes and explicit diffusion only, and
 * returns CA_ERR_ARG otherwise, before any launch.
 * ca_step that also keeps what its pullback needs: runs nsteps steps of Y in
 * place (as ca_step) and stores the state at every `every`-th step in the
 * handle (ceil(nsteps / every) states; ca_step_pullback adds `every` steps of
 * stage records, 7 states + 4 Laplacian sets each).  every = 0 keeps the
 * stage records of all nsteps steps instead (nsteps x (7 states + 4 Laplacian
 * sets) of device memory): the pullback then only reverses, with no
 * recomputation, and gives bitwise the same cotangents. */
ca_status ca_step_record(ca_handle *h, float *Y, double dt, int32_t nsteps, int32_t every, void *stream);
/* The pullback of the last ca_step_record: given the cotangent YbN of its
 * result, writes the cotangent of its input to Yb0 and, with vertical
 * diffusion and D0b != NULL, each column's D0 cotangent to D0b [ncol].  Each
 * checkpoint segment is re-run once with its stage states recorded, then
 * reversed.  The column parameters must be those of the recorded run. */
ca_status ca_step_pullback(ca_handle *h, const float *YbN, float *Yb0, float *D0b, void *stream);
/* ca_step_record on a copy of Y0 followed by ca_step_pullback (every as
 * there, 0 included).
 * The pullback of ca_step(h, Y0, dt, nsteps): given the cotangent YbN of the
 * state after nsteps steps from Y0, writes the cotangent of Y0 to Yb0 and,
 * with vertical diffusion and D0b != NULL, the cotangent of each column's D0
 * to D0b [ncol]. */
ca_status ca_step_adjoint(ca_handle *h, const float *Y0, const float *YbN, double dt, int32_t nsteps,
                          int32_t every, float *Yb0, float *D0b, void *stream);
/* New per-column Rayleigh coefficients and / or vertical diffusivities
 * (host arrays of ncol; NULL keeps), for handles created with them (the
 * reference's parameter fields; calibration changes them between runs). */
ca_status ca_set_column_params(ca_handle *h, const double *alpha_w_col, const double *D0_col);

#ifdef __cplusplus
}
#endif
#endif
