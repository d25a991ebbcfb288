// ca_step_adjoint_api.hip -- host side of the sphere step's reverse mode:
// ca_dss_vjp, ca_remaining_tendency_vjp, ca_step_record / ca_step_pullback,
// ca_step_adjoint and ca_set_column_params.  Included after
// ca_api.hip in the single translation unit (it uses the handle internals).
#include "ca_kernels.h"

namespace {

// planes of the S-sized reverse-mode scratch h->aw_buf
enum AdjBuf { AB_UB, AB_CB, AB_UTB1, AB_UTB2, AB_UTB3, AB_FEB, AB_FIB, AB_GB, AB_UBS, AB_Y, AB_YF, AB_COUNT };

ca_status adj_ensure(ca_handle *h, int64_t rec_steps, int64_t nck) {
    const DevGeom &g = h->g;
    ca_status s;
    if (!h->aw_buf) {
        if ((s = dalloc(h, &h->aw_buf, (size_t)AB_COUNT * h->S))) return s;
        if ((s = dalloc(h, &h->aw_L, 16 * (size_t)g.nC))) return s;
        HIPCHK(h, hipMemset(h->aw_buf, 0xff, (size_t)AB_COUNT * h->S * sizeof(float)));
        HIPCHK(h, hipMemset(h->aw_L, 0xff, 16 * (size_t)g.nC * sizeof(float)));
        const int nch = (int)cdiv(g.nlev, h->adj_chunk);
        if ((s = dalloc(h, &h->aw_d0, (size_t)nch * g.ncol))) return s;
        // per-column one-column views of the geometry (the Newton stage reverse)
        std::vector<DevGeom> v(g.ncol);
        for (int c = 0; c < g.ncol; ++c) {
            DevGeom cg = g;
            cg.nelem = 0;
            cg.ncol = 1;
            cg.nC = g.nlev;
            cg.nF = g.nlev + 1;
            for (const float **p : {&cg.Jh, &cg.gu11, &cg.gu12, &cg.gu22, &cg.gd11, &cg.gd12, &cg.gd22, &cg.f3n,
                                    &cg.f12a, &cg.f12b, &cg.iJh})
                *p += c;
            cg.basis = nullptr;
            if (cg.aw_col) cg.aw_col += c;
            if (cg.D0_col) cg.D0_col += c;
            cg.gtc = nullptr;
            cg.gtw = cg.gtM = nullptr;
            cg.ncl = 0;
            cg.dl_list = cg.dl_nE = cg.dl_col = nullptr;
            cg.dl_slot = nullptr;
            cg.dl_w = cg.dl_M = nullptr;
            v[c] = cg;
        }
        if ((s = upload(h, &h->cgeo, v))) return s;
    }
    // the records and checkpoints grow with the call's `every` and step
    // count; a larger request replaces the smaller buffer
    if (rec_steps > h->aw_rec_steps) {
        if ((s = dfree(h, h->aw_rec))) return s;
        h->aw_rec = nullptr;
        h->aw_rec_steps = 0;
        if ((s = dalloc(h, &h->aw_rec, (size_t)(rec_steps * rec_size(h))))) return s;
        HIPCHK(h, hipMemset(h->aw_rec, 0xff, (size_t)(rec_steps * rec_size(h)) * sizeof(float)));
        h->aw_rec_steps = rec_steps;
    }
    if (nck > h->aw_ck_n) {
        if ((s = dfree(h, h->aw_ck))) return s;
        h->aw_ck = nullptr;
        h->aw_ck_n = 0;
        if ((s = dalloc(h, &h->aw_ck, (size_t)(nck * h->S)))) return s;
        HIPCHK(h, hipMemset(h->aw_ck, 0xff, (size_t)(nck * h->S) * sizeof(float)));
        h->aw_ck_n = nck;
    }
    return CA_OK;
}

// Debug poisoning of the reverse-mode scratch (see poison_work); with
// records, also the per-step records, which every pullback rewrites.  The
// AB_YF plane (the last one) is left alone: it holds ca_step_adjoint's copy
// of its input across the ca_step_record it makes, whose poisoning would
// otherwise erase it; it is written before every read.
static_assert(AB_YF == AB_COUNT - 1, "AB_YF must be the last scratch plane");
ca_status poison_reverse(ca_handle *h, hipStream_t st, bool records) {
    if (!h->poison) return CA_OK;
    ca_status s;
    if ((s = poison_work(h, st))) return s;
    if ((s = fill_words(h, h->aw_buf, (int64_t)AB_YF * h->S, kNaNBits, st))) return s;
    if ((s = fill_words(h, h->aw_L, 16 * h->g.nC, kNaNBits, st))) return s;
    if (records && (s = fill_words(h, h->aw_rec, h->aw_rec_steps * rec_size(h), kNaNBits, st))) return s;
    return CA_OK;
}

// the Newton-stage reverse keeps a column's faces in one workgroup's LDS rows
// of at most 129 entries
constexpr int kMaxAdjLevels = 128;

float *abuf(ca_handle *h, int i) { return h->aw_buf + (int64_t)i * h->S; }

// DSS^T of a state-shaped buffer, src -> dst (filter: also the u_3 boundary filter)
ca_status dss_T_state(ca_handle *h, const float *src, float *dst, int filter, hipStream_t st) {
    const DevGeom &g = h->gT;
    const int ch = h->dss_chunk > 0 ? h->dss_chunk : g.nlev;
    if (h->dss_kind == 0 && g.ncl > 0)
        launch_dss_lds<4, 1, true>(g, src, dst, filter, ch, st);
    else
        k_dss_gather<4, 1, true><<<dss_grid(g, ch), 256, 0, st>>>(g, src, dst, filter, ch);
    return check_launch(h, "k_dss(transposed state)");
}

struct LinTerm {
    const float *x;
    double a, a2;
};

// out = (acc ? out : 0) + sum a x over `planes` (+ the Laplacian face terms
// fa, fb on the interior faces); with out2: out2 = sum a2 x over planes2
ca_status lincomb(ca_handle *h, float *out, std::initializer_list<LinTerm> terms, bool acc, unsigned planes,
                  hipStream_t st, float *out2 = nullptr, unsigned planes2 = 0, const float *fa = nullptr,
                  const float *fb = nullptr) {
    sadj::Lin L{};
    L.nx = 0;
    for (const auto &t : terms) {
        if (t.a == 0.0 && t.a2 == 0.0) continue;
        if (L.nx >= 6) return fail(h, CA_ERR_ARG, "lincomb: too many terms");
        L.x[L.nx] = t.x;
        L.a[L.nx] = (float)t.a;
        L.a2[L.nx] = (float)t.a2;
        ++L.nx;
    }
    L.acc = acc ? 1 : 0;
    L.planes = planes;
    L.out2 = out2;
    L.planes2 = planes2;
    L.fa = fa;
    L.fb = fb;
    auto al = [](const void *p) { return ((uintptr_t)p & 15) == 0; };
    bool v4 = h->g.ncol % 4 == 0 && al(out) && (!out2 || al(out2)) && (!fa || (al(fa) && al(fb)));
    for (int t = 0; t < L.nx; ++t) v4 = v4 && al(L.x[t]);
    if (v4)
        sadj::k_lincomb4<<<cdiv(cdiv(h->S, 4), 256), 256, 0, st>>>(h->g, out, L);
    else
        sadj::k_lincomb<<<cdiv(h->S, 256), 256, 0, st>>>(h->g, out, L);
    return check_launch(h, "k_lincomb");
}

constexpr unsigned PL_ALL = 0x1F, PL_FI = (1u << 0) | (1u << 3) | (1u << 4);  // rho, rho e_tot, u_3

ca_status fold_faces(ca_handle *h, float *Yb, const float *fa, const float *fb, hipStream_t st) {
    sadj::k_fold_faces<<<dim3(cdiv(h->g.ncol, 256), h->g.nlev - 1), 256, 0, st>>>(h->g, Yb, fa, fb);
    return check_launch(h, "k_fold_faces");
}

// Yb = (d T_E / d U)^T gb at U with the DSSed Laplacians L (hyperdiffusion);
// D0 cotangents are added to h->aw_d0.  gb: cotangent of the un-DSSed
// explicit tendency.
// With fold = false the u_3 face terms of the explicit and Laplacian
// reverses are left in the handle's fa / fb planes (face_a, face_b) for the
// caller to add.
ca_status explicit_vjp(ca_handle *h, const float *U, const float *L, const float *gb, float *Yb, hipStream_t st,
                       bool fold = true) {
    const DevGeom &g = h->g;
    const DevNum &n = h->n;
    const int ch = h->adj_chunk;
    const dim3 grid(cdiv(g.ncol, 256), cdiv(g.nlev, ch)), blk(256);
    float *Lb = h->aw_L, *Lb2 = h->aw_L + 4 * g.nC, *fa = h->aw_L + 8 * g.nC, *fb = h->aw_L + 12 * g.nC;
    const bool hy = n.hyper != 0, vd = n.vdiff != 0;
    if (hy && vd)
        sadj::k_explicit_vjp<true, true><<<grid, blk, 0, st>>>(g, n, U, L, gb, Yb, Lb, fa, fb, h->aw_d0, ch);
    else if (hy)
        sadj::k_explicit_vjp<true, false><<<grid, blk, 0, st>>>(g, n, U, L, gb, Yb, Lb, fa, fb, h->aw_d0, ch);
    else if (vd)
        sadj::k_explicit_vjp<false, true><<<grid, blk, 0, st>>>(g, n, U, L, gb, Yb, Lb, fa, fb, h->aw_d0, ch);
    else
        sadj::k_explicit_vjp<false, false><<<grid, blk, 0, st>>>(g, n, U, L, gb, Yb, Lb, fa, fb, h->aw_d0, ch);
    ca_status s = check_launch(h, "k_explicit_vjp");
    if (s) return s;
    if (!hy) return fold && g.nlev >= 2 ? fold_faces(h, Yb, fa, fb, st) : CA_OK;
    {
        const DevGeom &gT = h->gT;
        const int dch = h->dss_chunk > 0 ? h->dss_chunk : gT.nlev;
        if (h->dss_kind == 0 && gT.ncl > 0)
            launch_dss_lds<4, 0, false>(gT, Lb, Lb2, 0, dch, st);
        else
            k_dss_gather<4, 0, false><<<dss_grid(gT, dch), 256, 0, st>>>(gT, Lb, Lb2, 0, dch);
    }
    if ((s = check_launch(h, "k_dss(transposed laplacians)"))) return s;
    sadj::k_laplacians_vjp<<<dim3(cdiv(g.ncol, 256), g.nlev), 256, 0, st>>>(g, n, U, Lb2, Yb, fa, fb);
    if ((s = check_launch(h, "k_laplacians_vjp"))) return s;
    return fold && g.nlev >= 2 ? fold_faces(h, Yb, fa, fb, st) : CA_OK;
}

// the face planes of the explicit and Laplacian reverses' u_3 terms
const float *face_a(const ca_handle *h) { return h->aw_L + 8 * h->g.nC; }
const float *face_b(const ca_handle *h) { return h->aw_L + 12 * h->g.nC; }

// Utb = (d Newton_i / d U~)^T Ub at U~ = Ut (Un: the stage's recorded
// result), column by column, + cfi fib on rho, rho e_tot and u_3
// The cotangent of U is Ub - cfi fib (rho, rho e_tot, u_3) + the face terms
// fa, fb (if given): F_I = (U - U~) / dtg with cfi = -1 / dtg.
ca_status newton_vjp(ca_handle *h, const float *Ut, const float *Un, const float *Ub, float dtg, float *Utb,
                     const float *fib, float cfi, const float *fa, const float *fb, hipStream_t st) {
    const DevGeom &g = h->g;
    if (g.nlev > kMaxAdjLevels) return fail(h, CA_ERR_ARG, "newton_vjp: nlev > 128");
    if (g.nlev <= 32)
        sadj::k_newton_vjp<32, 33, 2><<<dim3(cdiv(g.ncol, 2)), 64, 0, st>>>(h->cgeo, h->n, g.nC, g.ncol, Ut, Un, Ub,
                                                                            dtg, Utb, fib, cfi, fa, fb);
    else if (g.nlev <= 64)
        sadj::k_newton_vjp<64, 65, 1><<<dim3(g.ncol), 64, 0, st>>>(h->cgeo, h->n, g.nC, g.ncol, Ut, Un, Ub, dtg, Utb,
                                                                   fib, cfi, fa, fb);
    else
        sadj::k_newton_vjp<128, 129, 1><<<dim3(g.ncol), 128, 0, st>>>(h->cgeo, h->n, g.nC, g.ncol, Ut, Un, Ub, dtg,
                                                                      Utb, fib, cfi, fa, fb);
    return check_launch(h, "k_newton_vjp");
}

ca_status zero_d0(ca_handle *h, hipStream_t st) {
    const int nch = (int)cdiv(h->g.nlev, h->adj_chunk);
    return fill_words(h, h->aw_d0, 2 * (int64_t)nch * h->g.ncol, 0u, st);
}

ca_status sum_d0(ca_handle *h, float *D0b, hipStream_t st) {
    const int nch = (int)cdiv(h->g.nlev, h->adj_chunk);
    sadj::k_d0_sum<<<cdiv(h->g.ncol, 256), 256, 0, st>>>(h->g.ncol, nch, h->aw_d0, D0b);
    return check_launch(h, "k_d0_sum");
}

// Reverse of one step (one_step with records rec): on entry AB_UB holds the
// cotangent of the step's result, on exit that of its start state.
ca_status reverse_step(ca_handle *h, float *rec, double dt, hipStream_t st) {
    ca_status s;
    float *ub = abuf(h, AB_UB), *Cb = abuf(h, AB_CB), *feb = abuf(h, AB_FEB), *fib = abuf(h, AB_FIB);
    float *gb = abuf(h, AB_GB), *Ubs = abuf(h, AB_UBS);
    float *Utb[4] = {nullptr, abuf(h, AB_UTB1), abuf(h, AB_UTB2), abuf(h, AB_UTB3)};
    const float *fa = face_a(h), *fb = face_b(h);
    // u' = DSS_f(C), C = u + dt sum_j b_j (F_E_j + F_I_j)
    if ((s = dss_T_state(h, ub, Cb, 1, st))) return s;
    for (int i = 3; i >= 1; --i) {
        const double dtg = dt * h->aI[i][i];
        // cotangents of F_E_i (feb) and F_I_i (fib): from C and the later stages' U~
        auto ae = [&](int l) { return l > i ? dt * h->aE[l][i] : 0.0; };
        auto ai = [&](int l) { return l > i ? dt * h->aI[l][i] : 0.0; };
        if ((s = lincomb(h, feb,
                         {{Cb, dt * h->b[i], dt * h->b[i]}, {Utb[1], ae(1), ai(1)}, {Utb[2], ae(2), ai(2)},
                          {Utb[3], ae(3), ai(3)}},
                         false, PL_ALL, st, fib, PL_FI)))
            return s;
        // F_E_i = DSS(T_E(U_i)) for i < 3; F_E_3 = T_E(U_3), which only the
        // end-of-step DSS averages (one_step)
        if (i < 3 && (s = dss_T_state(h, feb, gb, 0, st))) return s;
        if ((s = explicit_vjp(h, rec_U(h, rec, i), rec_L(h, rec, i), i < 3 ? gb : feb, Ubs, st, false))) return s;
        // U_i = Newton(U~_i); F_I_i = (U_i - U~_i) / dtg and the Laplacian
        // face terms are added to the U_i cotangent as the kernel loads it
        if ((s = newton_vjp(h, rec_Ut(h, rec, i), rec_U(h, rec, i), Ubs, (float)dtg, Utb[i], fib, (float)(-1.0 / dtg),
                            fa, fb, st)))
            return s;
    }
    // F_E_0 = DSS(T_E(u))
    if ((s = lincomb(h, feb,
                     {{Cb, dt * h->b[0], 0.0}, {Utb[1], dt * h->aE[1][0], 0.0}, {Utb[2], dt * h->aE[2][0], 0.0},
                      {Utb[3], dt * h->aE[3][0], 0.0}},
                     false, PL_ALL, st)))
        return s;
    if ((s = dss_T_state(h, feb, gb, 0, st))) return s;
    if ((s = explicit_vjp(h, rec_U(h, rec, 0), rec_L(h, rec, 0), gb, Ubs, st, false))) return s;
    // u enters C and every U~_i with weight 1
    return lincomb(h, ub, {{Cb, 1.0, 0.0}, {Utb[1], 1.0, 0.0}, {Utb[2], 1.0, 0.0}, {Utb[3], 1.0, 0.0}, {Ubs, 1.0, 0.0}},
                   false, PL_ALL, st, nullptr, 0, fa, fb);
}

}  // namespace

extern "C" {

ca_status ca_dss_vjp(ca_handle *h, float *Yb, void *stream) {
    if (!h || !Yb) return CA_ERR_ARG;
    if (h->g.nelem == 0) return CA_OK;
    hipStream_t st = (hipStream_t)stream;
    ca_status s;
    if ((s = adj_ensure(h, 0, 0))) return s;
    if ((s = poison_reverse(h, st, false))) return s;
    float *tmp = abuf(h, AB_GB);
    HIPCHK(h, hipMemcpyAsync(tmp, Yb, (size_t)h->S * sizeof(float), hipMemcpyDeviceToDevice, st));
    return dss_T_state(h, tmp, Yb, 0, st);
}

ca_status ca_remaining_tendency_vjp(ca_handle *h, float *Yb, const float *Ytb, const float *Y, float *D0b,
                                    void *stream) {
    if (!h || !Yb || !Ytb || !Y) return CA_ERR_ARG;
    if (!h->n.sphere) return fail(h, CA_ERR_ARG, "ca_remaining_tendency_vjp: sphere grids only (columns: ca_ensemble_adjoint)");
    if (h->n.sflux) return fail(h, CA_ERR_ARG, "ca_remaining_tendency_vjp: the reverse mode has no surface fluxes");
    if (h->n.idiff) return fail(h, CA_ERR_ARG, "ca_remaining_tendency_vjp: the reverse mode has no implicit diffusion");
    if (!h->vd_plain)
        return fail(h, CA_ERR_ARG, "ca_remaining_tendency_vjp: the reverse mode takes DecayWithHeightDiffusion with H = inf only");
    hipStream_t st = (hipStream_t)stream;
    const DevGeom &g = h->g;
    ca_status s;
    if ((s = adj_ensure(h, 0, 0))) return s;
    if ((s = poison_reverse(h, st, false))) return s;
    if (h->n.hyper) {  // the DSSed Laplacians at Y
        {
            Launch l(h, st, KID_LAPLACIANS);
            k_laplacians<<<dim3(cdiv(g.ncol, 256), cdiv(g.nlev, h->lap_chunk)), 256, 0, st>>>(g, h->n, Y, h->raw,
                                                                                       h->lap_chunk);
        }
        {
            Launch l(h, st, KID_DSS_LAPLACIAN);
            launch_dss<4, 0, false>(h, h->raw, h->L, 0, st);
        }
        if ((s = check_launch(h, "laplacians"))) return s;
    }
    if ((s = zero_d0(h, st))) return s;
    if ((s = explicit_vjp(h, Y, h->L, Ytb, Yb, st))) return s;
    if (D0b && h->n.vdiff) return sum_d0(h, D0b, st);
    return CA_OK;
}

ca_status ca_step_record(ca_handle *h, float *Y, double dt, int32_t nsteps, int32_t every, void *stream) {
    if (!h || !Y || nsteps < 1 || every < 0 || !(dt > 0)) return CA_ERR_ARG;
    if (!h->n.sphere) return fail(h, CA_ERR_ARG, "ca_step_record: sphere grids only (columns: ca_ensemble_adjoint)");
    if (h->g.nlev > kMaxAdjLevels) return fail(h, CA_ERR_ARG, "ca_step_record: the reverse mode supports nlev <= 128");
    if (h->n.sflux) return fail(h, CA_ERR_ARG, "ca_step_record: the reverse mode has no surface fluxes");
    if (h->n.idiff) return fail(h, CA_ERR_ARG, "ca_step_record: the reverse mode has no implicit diffusion");
    if (!h->vd_plain)
        return fail(h, CA_ERR_ARG, "ca_step_record: the reverse mode takes DecayWithHeightDiffusion with H = inf only");
    hipStream_t st = (hipStream_t)stream;
    const int64_t S = h->S;
    const bool all = every == 0;  // the stage records of every step, no checkpoints
    const int nseg = all ? 0 : (nsteps + every - 1) / every;
    ca_status s;
    h->rec_valid = false;
    if ((s = adj_ensure(h, all ? nsteps : every, nseg))) return s;
    if ((s = poison_reverse(h, st, true))) return s;
    if (h->poison && h->aw_ck && (s = fill_words(h, h->aw_ck, h->aw_ck_n * S, kNaNBits, st))) return s;
    // the step's initial dss! and filter, then a checkpoint at every segment start
    HIPCHK(h, hipMemcpyAsync(h->U, Y, (size_t)S * sizeof(float), hipMemcpyDeviceToDevice, st));
    if ((s = dss_gather_state(h, h->U, Y, 1, st, KID_DSS_STATE))) return s;
    const int64_t RS = rec_size(h);
    for (int t = 0; t < nsteps; ++t) {
        if (!all && t % every == 0)
            HIPCHK(h, hipMemcpyAsync(h->aw_ck + (int64_t)(t / every) * S, Y, (size_t)S * sizeof(float),
                                     hipMemcpyDeviceToDevice, st));
        if ((s = one_step(h, Y, dt, st, all ? h->aw_rec + (int64_t)t * RS : nullptr))) return s;
    }
    h->rec_valid = true;
    h->rec_dt = dt;
    h->rec_nsteps = nsteps;
    h->rec_every = every;
    return CA_OK;
}

ca_status ca_step_pullback(ca_handle *h, const float *YbN, float *Yb0, float *D0b, void *stream) {
    if (!h || !YbN || !Yb0) return CA_ERR_ARG;
    if (!h->rec_valid) return fail(h, CA_ERR_ARG, "ca_step_pullback: no trajectory recorded (ca_step_record)");
    hipStream_t st = (hipStream_t)stream;
    const int64_t S = h->S;
    const int nsteps = h->rec_nsteps, every = h->rec_every;
    const double dt = h->rec_dt;
    const bool all = every == 0;
    const int nseg = all ? 0 : (nsteps + every - 1) / every;
    ca_status s;
    // with every step's records kept by the forward, they are the input here
    if ((s = poison_reverse(h, st, !all))) return s;
    float *Yw = abuf(h, AB_Y);
    if ((s = zero_d0(h, st))) return s;
    HIPCHK(h, hipMemcpyAsync(abuf(h, AB_UB), YbN, (size_t)S * sizeof(float), hipMemcpyDeviceToDevice, st));
    const int64_t RS = rec_size(h);
    if (all)
        for (int j = nsteps - 1; j >= 0; --j)
            if ((s = reverse_step(h, h->aw_rec + (int64_t)j * RS, dt, st))) return s;
    // per segment (last first): re-run it from its checkpoint with stage records, then reverse its steps
    for (int sg = nseg - 1; sg >= 0; --sg) {
        const int len = std::min(every, nsteps - sg * every);
        HIPCHK(h, hipMemcpyAsync(Yw, h->aw_ck + (int64_t)sg * S, (size_t)S * sizeof(float), hipMemcpyDeviceToDevice,
                                 st));
        for (int j = 0; j < len; ++j)
            if ((s = one_step(h, Yw, dt, st, h->aw_rec + j * RS))) return s;
        for (int j = len - 1; j >= 0; --j)
            if ((s = reverse_step(h, h->aw_rec + j * RS, dt, st))) return s;
    }
    // the initial dss! and filter
    if ((s = dss_T_state(h, abuf(h, AB_UB), Yb0, 1, st))) return s;
    if (D0b && h->n.vdiff) return sum_d0(h, D0b, st);
    return CA_OK;
}

ca_status ca_step_adjoint(ca_handle *h, const float *Y0, const float *YbN, double dt, int32_t nsteps, int32_t every,
                          float *Yb0, float *D0b, void *stream) {
    if (!h || !Y0 || !YbN || !Yb0 || nsteps < 1 || every < 0 || !(dt > 0)) return CA_ERR_ARG;
    if (!h->n.sphere) return fail(h, CA_ERR_ARG, "ca_step_adjoint: sphere grids only (columns: ca_ensemble_adjoint)");
    if (h->g.nlev > kMaxAdjLevels) return fail(h, CA_ERR_ARG, "ca_step_adjoint: the reverse mode supports nlev <= 128");
    if (h->n.sflux) return fail(h, CA_ERR_ARG, "ca_step_adjoint: the reverse mode has no surface fluxes");
    if (h->n.idiff) return fail(h, CA_ERR_ARG, "ca_step_adjoint: the reverse mode has no implicit diffusion");
    if (!h->vd_plain)
        return fail(h, CA_ERR_ARG, "ca_step_adjoint: the reverse mode takes DecayWithHeightDiffusion with H = inf only");
    hipStream_t st = (hipStream_t)stream;
    ca_status s;
    if ((s = adj_ensure(h, every == 0 ? nsteps : every, every == 0 ? 0 : (nsteps + every - 1) / every))) return s;
    float *Yf = abuf(h, AB_YF);
    HIPCHK(h, hipMemcpyAsync(Yf, Y0, (size_t)h->S * sizeof(float), hipMemcpyDeviceToDevice, st));
    if ((s = ca_step_record(h, Yf, dt, nsteps, every, stream))) return s;
    return ca_step_pullback(h, YbN, Yb0, D0b, stream);
}

ca_status ca_set_column_params(ca_handle *h, const double *alpha_w_col, const double *D0_col) {
    if (!h) return CA_ERR_ARG;
    const int nc = h->g.ncol;
    if (alpha_w_col) {
        if (!h->g.aw_col) return fail(h, CA_ERR_ARG, "ca_set_column_params: created without alpha_w_col");
        std::vector<float> v(alpha_w_col, alpha_w_col + nc);
        HIPCHK(h, hipMemcpy((void *)h->g.aw_col, v.data(), nc * sizeof(float), hipMemcpyHostToDevice));
    }
    if (D0_col) {
        if (!h->g.D0_col) return fail(h, CA_ERR_ARG, "ca_set_column_params: created without vertical diffusion");
        std::vector<float> v(D0_col, D0_col + nc);
        HIPCHK(h, hipMemcpy((void *)h->g.D0_col, v.data(), nc * sizeof(float), hipMemcpyHostToDevice));
    }
    h->rec_valid = false;
    return CA_OK;
}

}  // extern "C"
