// ca_ensemble_api.hip -- host side of the column ensemble (ca_ensemble_*).
// Included after ca_api.hip in the single translation unit: it shares the
// level-table and ARS343 helpers.
#include <mutex>

#include "ca_kernels.h"

// ca_ensemble_adjoint_noalloc's scratch: 16 Mi floats (64 MiB) per GPU
#define CA_NOALLOC_SCRATCH_FLOATS (int64_t(1) << 24)
#define CA_MAX_DEVICES 64

struct ca_ensemble {
    int device = 0;
    int M = 0;
    DevNum n{};
    EnsTab tab{};
    std::vector<void *> allocs;
    std::vector<int64_t> off;
    std::vector<int> nlev;
    int64_t S = 0;
    DevGeom *geoms = nullptr;
    int *order[2] = {nullptr, nullptr};  // 0: as given, 1: longest first
    int64_t *doff = nullptr;
    int lanes = 64;
    int order_kind = 1;
    int nf = CA_ENS_NF;  // LDS face rows per member: 33, 65 or 129, from the tallest member
    std::string err;
    // one captured launch, replayed while (Y, dt, nsteps, schedule, stream) match
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    float *gY = nullptr;
    double gdt = 0.0;
    int gn = 0, glanes = 0, gorder = -1;
    hipStream_t gst = nullptr;
    // per-member column constants (13 floats each) and their host copy
    float *cols = nullptr;
    std::vector<float> hcols;
    // adjoint scratch: checkpoints and segment states
    float *ckpt = nullptr, *segb = nullptr;
    int ck_nsteps = -1, ck_every = -1;
    // this device's copy of the file-scope scratch array (ca_ensemble_adjoint_noalloc)
    float *noalloc_scratch = nullptr;
};

namespace {

ca_status efail(ca_ensemble *e, ca_status s, const std::string &msg) {
    if (e) e->err = msg;
    return s;
}

#define EHIPCHK(e, call)                                                                  \
    do {                                                                                  \
        hipError_t e_ = (call);                                                           \
        if (e_ != hipSuccess)                                                             \
            return efail(e, CA_ERR_DEVICE, std::string(#call ": ") + hipGetErrorString(e_)); \
    } while (0)

template <class T>
ca_status eupload(ca_ensemble *e, T **dst, const std::vector<T> &src) {
    void *q = nullptr;
    EHIPCHK(e, hipMalloc(&q, (src.empty() ? 1 : src.size()) * sizeof(T) + 16));
    e->allocs.push_back(q);
    if (!src.empty()) EHIPCHK(e, hipMemcpy(q, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice));
    *dst = (T *)q;
    return CA_OK;
}

void ens_drop_graph(ca_ensemble *e) {
    if (e->exec) (void)hipGraphExecDestroy(e->exec);
    if (e->graph) (void)hipGraphDestroy(e->graph);
    e->exec = nullptr;
    e->graph = nullptr;
}

ca_status ens_launch(ca_ensemble *e, float *Y, float dt, int nsteps, hipStream_t st) {
    const int *ord = e->order[e->order_kind];
    const dim3 grid(e->M);
#define CA_ENS_GO(LPM, NF) k_ensemble<LPM, NF><<<grid, LPM, 0, st>>>(e->geoms, e->n, ord, e->doff, Y, e->tab, dt, nsteps)
    if (e->nf == 33) {  // at most 32 levels: 256 lanes would idle, run 128
        if (e->lanes == 64) CA_ENS_GO(64, 33);
        else CA_ENS_GO(128, 33);
    } else if (e->nf == 65) {
        if (e->lanes == 64) CA_ENS_GO(64, 65);
        else if (e->lanes == 128) CA_ENS_GO(128, 65);
        else CA_ENS_GO(256, 65);
    } else {
        if (e->lanes == 64) CA_ENS_GO(64, 129);
        else if (e->lanes == 128) CA_ENS_GO(128, 129);
        else CA_ENS_GO(256, 129);
    }
#undef CA_ENS_GO
    EHIPCHK(e, hipGetLastError());
    return CA_OK;
}

ca_status ens_create_impl(ca_ensemble *e, const ca_params *par, const ca_config *cfg, int M, const ca_member *mem) {
    EHIPCHK(e, hipGetDevice(&e->device));
    e->M = M;
    // numerics shared by all members (the column profile: no hyperdiffusion)
    DevNum &n = e->n;
    n.R_d = (float)par->R_d;
    n.cp_d = (float)par->cp_d;
    n.cv_d = (float)par->cv_d;
    n.T_0 = (float)par->T_0;
    n.p0inv = (float)(1.0 / par->p_ref_theta);
    n.kex = (float)(par->R_d / par->cp_d);
    n.kap = (float)(par->R_d / par->cv_d);
    n.grav = (float)par->grav;
    n.Tsr = (float)par->T_surf_ref;
    n.Tmr = (float)par->T_min_ref;
    n.Tmin = (float)par->T_min_sgs;
    n.dR_d = par->R_d;
    n.dcp_d = par->cp_d;
    n.dcv_d = par->cv_d;
    n.dicv_d = 1.0 / par->cv_d;
    n.dT_0 = par->T_0;
    n.dp0inv = 1.0 / par->p_ref_theta;
    n.dkex = par->R_d / par->cp_d;
    n.dkex_r = n.dkex - 2.0 / 7.0;
    if (std::fabs(n.dkex_r) > 1e-6)
        return efail(e, CA_ERR_ARG, "ca_ensemble_create: R_d / cp_d must be 2/7 to within 1e-6 (dry air; the float64 "
                                    "Exner function refines q^(2/7))");
    n.dTsr = par->T_surf_ref;
    n.dTmr = par->T_min_ref;
    n.dTmin = par->T_min_sgs;
    n.hyper = 0;
    n.upw = cfg->upwinding;
    n.rayleigh = cfg->rayleigh;
    n.alpha_uh = (float)cfg->alpha_uh;
    n.alpha_w = (float)cfg->alpha_w;
    n.vdiff = cfg->vert_diff;
    n.dt = (float)cfg->dt;
    n.sphere = 0;
    // ARS343 tableau (the same solve for gamma as ca_step's)
    {
        ca_handle tmp;
        ars343(&tmp);
        for (int i = 0; i < 4; ++i) {
            for (int j = 0; j < 4; ++j) {
                e->tab.aE[i][j] = (float)tmp.aE[i][j];
                e->tab.aI[i][j] = (float)tmp.aI[i][j];
            }
            e->tab.b[i] = (float)tmp.b[i];
        }
    }
    // per-member level tables, packed; per-member column constants
    std::vector<float> lc, lf, col;
    std::vector<double> dphid, phicd;
    std::vector<size_t> olc(M), olf(M), od(M), oc(M), ocol(M);
    e->off.resize(M);
    e->nlev.resize(M);
    int64_t S = 0;
    for (int m = 0; m < M; ++m) {
        const int nl = mem[m].nlev;
        if (nl < 3 || nl > CA_ENS_MAXL || !mem[m].z_face)
            return efail(e, CA_ERR_ARG, "member " + std::to_string(m) + ": nlev must be in [3, 128] with z_face");
        for (int f = 0; f < nl; ++f)
            if (!(mem[m].z_face[f + 1] > mem[m].z_face[f]))
                return efail(e, CA_ERR_ARG, "member " + std::to_string(m) + ": z_face must increase");
        const LevelTables T = level_tables(nl, mem[m].z_face, 0.0, par->grav, cfg->zd_rayleigh);
        olc[m] = lc.size();
        olf[m] = lf.size();
        od[m] = dphid.size();
        oc[m] = phicd.size();
        lc.insert(lc.end(), T.lc.begin(), T.lc.end());
        lf.insert(lf.end(), T.lf.begin(), T.lf.end());
        dphid.insert(dphid.end(), T.dphid.begin(), T.dphid.end());
        phicd.insert(phicd.end(), T.phicd.begin(), T.phicd.end());
        // Jh, gu11, gu12, gu22, gd11, gd12, gd22, f3n, f12a, f12b, iJh, alpha_w, D0
        ocol[m] = col.size();
        const float cv[13] = {1, 1, 0, 1, 1, 0, 1, (float)mem[m].f3, 0, 0, 1, (float)mem[m].alpha_w, (float)mem[m].D0};
        col.insert(col.end(), cv, cv + 13);
        e->off[m] = S;
        e->nlev[m] = nl;
        S += 5 * (int64_t)nl + 1;
    }
    e->S = S;
    const int maxl = *std::max_element(e->nlev.begin(), e->nlev.end());
    e->nf = maxl <= 32 ? 33 : (maxl <= 64 ? 65 : CA_ENS_NF);
    float *dlc = nullptr, *dlf = nullptr, *dcol = nullptr;
    double *ddp = nullptr, *dpc = nullptr;
    ca_status s;
    if ((s = eupload(e, &dlc, lc))) return s;
    if ((s = eupload(e, &dlf, lf))) return s;
    if ((s = eupload(e, &ddp, dphid))) return s;
    if ((s = eupload(e, &dpc, phicd))) return s;
    if ((s = eupload(e, &dcol, col))) return s;
    e->cols = dcol;
    e->hcols = col;
    std::vector<DevGeom> geoms(M);
    for (int m = 0; m < M; ++m) {
        DevGeom &g = geoms[m];
        std::memset(&g, 0, sizeof(g));
        const int nl = mem[m].nlev;
        g.nelem = 0;
        g.ncol = 1;
        g.nlev = nl;
        g.nC = nl;
        g.nF = nl + 1;
        point_level_tables(g, nl, dlc + olc[m], dlf + olf[m], ddp + od[m], dpc + oc[m]);
        const float *c = dcol + ocol[m];
        g.Jh = c + 0;
        g.gu11 = c + 1;
        g.gu12 = c + 2;
        g.gu22 = c + 3;
        g.gd11 = c + 4;
        g.gd12 = c + 5;
        g.gd22 = c + 6;
        g.f3n = c + 7;
        g.f12a = c + 8;
        g.f12b = c + 9;
        g.iJh = c + 10;
        g.aw_col = c + 11;
        g.D0_col = c + 12;
    }
    if ((s = eupload(e, &e->geoms, geoms))) return s;
    std::vector<int> ord0(M), ord1(M);
    for (int m = 0; m < M; ++m) ord0[m] = ord1[m] = m;
    std::stable_sort(ord1.begin(), ord1.end(), [&](int a, int b) { return mem[a].nlev > mem[b].nlev; });
    if ((s = eupload(e, &e->order[0], ord0))) return s;
    if ((s = eupload(e, &e->order[1], ord1))) return s;
    if ((s = eupload(e, &e->doff, e->off))) return s;
    return CA_OK;
}

// Scratch of ca_ensemble_adjoint_noalloc: a file-scope device array, one
// copy per GPU, held until the process ends.  Host code never passes its
// name to a kernel (on the host the name is a stand-in, not a device
// address): ca_ensemble_create resolves this device's address once with
// hipGetSymbolAddress and every use goes through that pointer.
static __device__ __align__(256) float scratch[CA_NOALLOC_SCRATCH_FLOATS];

std::mutex g_scratch_mu;
float *g_scratch_ptr[CA_MAX_DEVICES] = {};

// this device's address of `scratch`, resolved once per device with the
// device current
ca_status resolve_noalloc_scratch(ca_ensemble *e) {
    int dev = 0;
    EHIPCHK(e, hipGetDevice(&dev));
    if (dev < 0 || dev >= CA_MAX_DEVICES) return efail(e, CA_ERR_DEVICE, "device ordinal out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        EHIPCHK(e, hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return efail(e, CA_ERR_DEVICE, "hipGetSymbolAddress(scratch) returned a null address");
        g_scratch_ptr[dev] = (float *)p;
    }
    e->noalloc_scratch = g_scratch_ptr[dev];
    return CA_OK;
}

// the reverse launch shared by ca_ensemble_adjoint and its no-allocation
// variant: ck holds (ceil(nsteps / every) + 1) * S floats, sg every * S
ca_status ens_adjoint_launch(ca_ensemble *e, const float *Y0, const float *YbN, double dt, int nsteps, int every,
                             float *Yb0, float *D0b, float *awb, float *ck, float *sg, hipStream_t st) {
    const int *ord = e->order[e->order_kind];
    const dim3 grid(e->M);
    // lanes: 64 for members of at most 32 levels; otherwise 64 or 128 (the
    // 256-lane forward schedule runs its adjoint with 128)
#define CA_ADJ_GO(LPM, NF)                                                                                       \
    k_ensemble_adjoint<LPM, NF><<<grid, LPM, 0, st>>>(e->geoms, e->n, ord, e->doff, Y0, YbN, Yb0, D0b, awb, ck, sg, \
                                                      e->tab, (float)dt, nsteps, every)
    if (e->nf == 33) CA_ADJ_GO(64, 33);
    else if (e->nf == 65 && e->lanes == 64) CA_ADJ_GO(64, 65);
    else if (e->nf == 65) CA_ADJ_GO(128, 65);
    else if (e->lanes == 64) CA_ADJ_GO(64, 129);
    else CA_ADJ_GO(128, 129);
#undef CA_ADJ_GO
    EHIPCHK(e, hipGetLastError());
    return CA_OK;
}

}  // namespace

extern "C" {

ca_status ca_ensemble_create(const ca_params *params, const ca_config *config, int32_t nmembers,
                             const ca_member *members, ca_ensemble **out) {
    if (!params || !config || !members || !out || nmembers < 1) return CA_ERR_ARG;
    *out = nullptr;
    if (config->upwinding != 0 && config->upwinding != 1 && config->upwinding != 3) return CA_ERR_ARG;
    if (config->sfc_flux) {
        g_create_error = "ca_ensemble_create: the column ensembles have no surface fluxes";
        return CA_ERR_ARG;
    }
    if (config->implicit_diffusion) {
        g_create_error = "ca_ensemble_create: the column ensembles have no implicit diffusion";
        return CA_ERR_ARG;
    }
    if (config->vert_diff && (config->vert_diff_model != 0 ||
                              (config->vert_diff_H > 0 && std::isfinite(config->vert_diff_H)))) {
        g_create_error = "ca_ensemble_create: the column ensembles take DecayWithHeightDiffusion with H = inf only";
        return CA_ERR_ARG;
    }
    ca_ensemble *e = new ca_ensemble();
    ca_status s = ens_create_impl(e, params, config, nmembers, members);
    if (s == CA_OK) s = resolve_noalloc_scratch(e);
    if (s != CA_OK) {
        g_create_error = e->err.empty() ? "ca_ensemble_create: invalid argument" : e->err;
        ca_ensemble_destroy(e);
        return s;
    }
    *out = e;
    return CA_OK;
}

ca_status ca_ensemble_destroy(ca_ensemble *e) {
    if (!e) return CA_ERR_ARG;
    ens_drop_graph(e);
    for (void *p : {(void *)e->ckpt, (void *)e->segb})
        if (p) (void)hipFree(p);
    for (void *p : e->allocs) (void)hipFree(p);
    delete e;
    return CA_OK;
}

int64_t ca_ensemble_state_size(const ca_ensemble *e) { return e ? e->S : 0; }

int64_t ca_ensemble_member_offset(const ca_ensemble *e, int32_t m) {
    if (!e || m < 0 || m >= e->M) return -1;
    return e->off[m];
}

const char *ca_ensemble_last_error(const ca_ensemble *e) { return e ? e->err.c_str() : g_create_error.c_str(); }

ca_status ca_ensemble_step(ca_ensemble *e, float *Y, double dt, int32_t nsteps, void *stream) {
    if (!e || !Y || nsteps < 0 || !(dt > 0)) return CA_ERR_ARG;
    if (nsteps == 0) return CA_OK;
    return ens_launch(e, Y, (float)dt, nsteps, (hipStream_t)stream);
}

ca_status ca_ensemble_step_graph(ca_ensemble *e, float *Y, double dt, int32_t nsteps, void *stream) {
    if (!e || !Y || nsteps < 1 || !(dt > 0) || !stream) return CA_ERR_ARG;
    hipStream_t st = (hipStream_t)stream;
    if (!(e->exec && e->gY == Y && e->gdt == dt && e->gn == nsteps && e->glanes == e->lanes &&
          e->gorder == e->order_kind && e->gst == st)) {
        ens_drop_graph(e);
        EHIPCHK(e, hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        const ca_status s = ens_launch(e, Y, (float)dt, nsteps, st);
        hipGraph_t graph = nullptr;
        const hipError_t ee = hipStreamEndCapture(st, &graph);
        if (s != CA_OK) {
            if (graph) (void)hipGraphDestroy(graph);
            return s;
        }
        if (ee != hipSuccess) return efail(e, CA_ERR_DEVICE, std::string("hipStreamEndCapture: ") + hipGetErrorString(ee));
        e->graph = graph;
        EHIPCHK(e, hipGraphInstantiate(&e->exec, e->graph, nullptr, nullptr, 0));
        e->gY = Y;
        e->gdt = dt;
        e->gn = nsteps;
        e->glanes = e->lanes;
        e->gorder = e->order_kind;
        e->gst = st;
    }
    EHIPCHK(e, hipGraphLaunch(e->exec, st));
    return CA_OK;
}

ca_status ca_ensemble_set_tuning(ca_ensemble *e, int32_t knob, int32_t value) {
    if (!e) return CA_ERR_ARG;
    if (knob == CA_ENS_LANES) {
        if (value != 64 && value != 128 && value != 256) return efail(e, CA_ERR_ARG, "lanes per member: 64, 128 or 256");
        e->lanes = value;
        return CA_OK;
    }
    if (knob == CA_ENS_ORDER) {
        if (value != 0 && value != 1) return efail(e, CA_ERR_ARG, "member order: 0 (as given) or 1 (longest first)");
        e->order_kind = value;
        return CA_OK;
    }
    return efail(e, CA_ERR_ARG, "unknown ensemble tuning knob");
}

ca_status ca_ensemble_get_tuning(const ca_ensemble *e, int32_t knob, int32_t *value) {
    if (!e || !value) return CA_ERR_ARG;
    if (knob == CA_ENS_LANES) *value = e->lanes;
    else if (knob == CA_ENS_ORDER) *value = e->order_kind;
    else return CA_ERR_ARG;
    return CA_OK;
}

ca_status ca_ensemble_autotune(ca_ensemble *e, const float *Y, double dt, int32_t nsteps, void *stream,
                               double *times_ms) {
    if (!e || !Y || nsteps < 1 || !(dt > 0)) return CA_ERR_ARG;
    hipStream_t st = (hipStream_t)stream;
    float *work = nullptr;
    EHIPCHK(e, hipMalloc(&work, (size_t)e->S * sizeof(float)));
    hipEvent_t a = nullptr, b = nullptr;
    ca_status s = CA_OK;
    const int lanes[3] = {64, 128, 256};
    double best = 1e300;
    int bl = e->lanes, bo = e->order_kind;
    if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) s = efail(e, CA_ERR_DEVICE, "hipEventCreate");
    for (int li = 0; li < 3 && s == CA_OK; ++li)
        for (int o = 0; o < 2 && s == CA_OK; ++o) {
            e->lanes = lanes[li];
            e->order_kind = o;
            double ms_best = 1e300;
            for (int rep = 0; rep < 2 && s == CA_OK; ++rep) {  // first repetition warms up
                if (hipMemcpyAsync(work, Y, (size_t)e->S * sizeof(float), hipMemcpyDeviceToDevice, st) != hipSuccess ||
                    hipEventRecord(a, st) != hipSuccess) {
                    s = efail(e, CA_ERR_DEVICE, "autotune: copy / event");
                    break;
                }
                if ((s = ens_launch(e, work, (float)dt, nsteps, st))) break;
                float ms = 0.0f;
                if (hipEventRecord(b, st) != hipSuccess || hipEventSynchronize(b) != hipSuccess ||
                    hipEventElapsedTime(&ms, a, b) != hipSuccess) {
                    s = efail(e, CA_ERR_DEVICE, "autotune: timing");
                    break;
                }
                ms_best = std::min(ms_best, (double)ms);
            }
            if (times_ms) times_ms[2 * li + o] = ms_best;
            if (ms_best < best) {
                best = ms_best;
                bl = lanes[li];
                bo = o;
            }
        }
    e->lanes = bl;
    e->order_kind = bo;
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
    (void)hipFree(work);
    return s;
}

ca_status ca_ensemble_set_params(ca_ensemble *e, const double *alpha_w, const double *D0) {
    if (!e) return CA_ERR_ARG;
    for (int m = 0; m < e->M; ++m) {
        if (alpha_w) e->hcols[13 * (size_t)m + 11] = (float)alpha_w[m];
        if (D0) e->hcols[13 * (size_t)m + 12] = (float)D0[m];
    }
    EHIPCHK(e, hipMemcpy(e->cols, e->hcols.data(), e->hcols.size() * sizeof(float), hipMemcpyHostToDevice));
    ens_drop_graph(e);
    return CA_OK;
}

ca_status ca_ensemble_adjoint(ca_ensemble *e, const float *Y0, const float *YbN, double dt, int32_t nsteps,
                              int32_t every, float *Yb0, float *D0b, float *awb, void *stream) {
    if (!e || !Y0 || !YbN || !Yb0 || !D0b || !awb || nsteps < 1 || every < 1 || !(dt > 0)) return CA_ERR_ARG;
    if (e->ck_nsteps != nsteps || e->ck_every != every) {
        for (float **p : {&e->ckpt, &e->segb})
            if (*p) {
                (void)hipFree(*p);
                *p = nullptr;
            }
        e->ck_nsteps = e->ck_every = -1;
        const int64_t nseg = (nsteps + every - 1) / every;
        EHIPCHK(e, hipMalloc(&e->ckpt, (size_t)((nseg + 1) * e->S) * sizeof(float)));
        EHIPCHK(e, hipMalloc(&e->segb, (size_t)(every * e->S) * sizeof(float)));
        e->ck_nsteps = nsteps;
        e->ck_every = every;
    }
    return ens_adjoint_launch(e, Y0, YbN, dt, nsteps, every, Yb0, D0b, awb, e->ckpt, e->segb, (hipStream_t)stream);
}

ca_status ca_ensemble_adjoint_noalloc(ca_ensemble *e, const float *Y0, const float *YbN, double dt, int32_t nsteps,
                                      int32_t every, float *Yb0, float *D0b, float *awb, void *stream) {
    if (!e || !Y0 || !YbN || !Yb0 || !D0b || !awb || nsteps < 1 || every < 1 || !(dt > 0)) return CA_ERR_ARG;
    if (!e->noalloc_scratch)
        return efail(e, CA_ERR_DEVICE, "ca_ensemble_adjoint_noalloc: the scratch array of this device is not resolved");
    const int64_t nseg = (nsteps + every - 1) / every;
    const int64_t need = (nseg + 1 + every) * e->S;
    if (need > CA_NOALLOC_SCRATCH_FLOATS)
        return efail(e, CA_ERR_ARG,
                     "ca_ensemble_adjoint_noalloc: needs (ceil(nsteps / every) + 1 + every) * state size = " +
                         std::to_string(need) + " floats of scratch, above the maximum of " +
                         std::to_string((long long)CA_NOALLOC_SCRATCH_FLOATS) +
                         " (use ca_ensemble_adjoint, or change every)");
    float *ck = e->noalloc_scratch;
    float *sg = ck + (nseg + 1) * e->S;
    return ens_adjoint_launch(e, Y0, YbN, dt, nsteps, every, Yb0, D0b, awb, ck, sg, (hipStream_t)stream);
}

}  // extern "C"
