// ca_api.hip -- host side of the C ABI declared in include/cadycore.h:
// handle life cycle, geometry upload, the hook-level entry points and the
// fused ARS343 step driver.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/cadycore.h"
#include "ca_kernels.h"
#include "ca_dss_table.h"


// message of the last failed ca_create (no handle exists to hold it)
static std::string g_create_error = "no error";

namespace {

enum KernelId {
    KID_LAPLACIANS = 0,
    KID_EXPLICIT,
    KID_DSS_TENDENCY,
    KID_DSS_LAPLACIAN,
    KID_NEWTON_STAGE,
    KID_FINAL,
    KID_DSS_STATE,
    KID_HOOKS,
    KID_EXPLICIT_FINAL,  // the last explicit tendency with the final combination
    KID_COUNT
};
const char *kKernelNames[KID_COUNT] = {"k_laplacians",   "k_explicit",   "k_dss(tendency)", "k_dss(laplacians)",
                                       "k_newton_stage", "k_final",      "k_dss(state)",    "hook kernels",
                                       "k_explicit_final"};

struct ProfRec {
    int id;
    hipEvent_t a, b;
};

}  // namespace

struct ca_handle {
    int device = 0;
    DevGeom g{};
    DevNum n{};
    int64_t S = 0;
    std::vector<void *> allocs;
    float *pre = nullptr;  // K, T, p, h planes
    float *L = nullptr;    // 4 Laplacian planes
    float *scr = nullptr;  // column scratch: 2 face planes + 6 centre planes
    float *FE[4] = {nullptr, nullptr, nullptr, nullptr};
    float *FI[4] = {nullptr, nullptr, nullptr, nullptr};
    float *U = nullptr;
    float *raw = nullptr;   // un-DSSed tendency / Laplacians (gather-DSS source)
    float *bands = nullptr;
    Bands B{};
    int *flag = nullptr;
    std::string err;
    bool prof = false;
    std::vector<ProfRec> recs;
    double aE[4][4], aI[4][4], b[4];
    int newton_variant = 1;
    bool vd_plain = true;  // no vertical diffusion, or DecayWithHeightDiffusion with H = inf (the reverse mode's)
    // ca_step_graph: one step's launches captured for replay, with what they bake in
    hipGraphExec_t gexec = nullptr;
    float *g_Y = nullptr;
    double g_dt = 0.0;
    hipStream_t g_st = nullptr;
    int explicit_chunk = 8;
    int final_chunk = 8;  // the last explicit launch's (the step's final combination)
    int dss_chunk = 8;
    int lap_chunk = 4;  // levels per work item of k_laplacians
    int dss_kind = 0;  // 0: LDS-staged per element cluster, 1: direct gather
    // reverse mode (ca_step_adjoint and the *_vjp hooks)
    DevGeom gT{};                 // g with the transposed DSS tables
    DevGeom *cgeo = nullptr;      // per-column one-column views of g (Newton stage reverse)
    int adj_chunk = 8;            // levels per work item of k_explicit_vjp
    float *aw_buf = nullptr;      // 13 S-sized cotangent / scratch planes
    float *aw_L = nullptr;        // 4 Laplacian-shaped planes: Lb, Lb2, fa | fb
    double *aw_d0 = nullptr;      // D0 cotangent partials [chunks x ncol]
    float *aw_rec = nullptr;      // per-step records of one checkpoint segment
    int64_t aw_rec_steps = 0;
    float *aw_ck = nullptr;       // checkpoints
    int64_t aw_ck_n = 0;
    // the trajectory ca_step_record last stored (for ca_step_pullback)
    bool rec_valid = false;
    double rec_dt = 0.0;
    int rec_nsteps = 0, rec_every = 0;
    // debug: fill the scratch no call carries over with NaN at every entry
    // point (the counterpart of fill_with_nans!, utils/debug_utils.jl:67-70)
    bool poison = false;
    // native-layout staging of the FieldVector hooks (ca_*_fv) for arguments
    // in another layout: allocated by the first such call, kept until destroy
    float *fv_a = nullptr, *fv_b = nullptr;
    // implicit vertical diffusion: the diffusion blocks and the solve's
    // column scratch (9 centre + 7 face planes); F_I holds all five
    // components then (fi_size = S), otherwise rho, rho e_tot and u_3
    DiffBands E{};
    float *idw = nullptr;
    int64_t fi_size = 0;
};

// 32-bit pattern fill of n words (NaN poisoning and zeroing; a kernel, so that
// a caller capturing the stream into a hipGraph records no memset node)
__global__ void k_fill_u32(uint32_t *p, int64_t n, uint32_t v) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

namespace {

ca_status fail(ca_handle *h, ca_status s, const std::string &msg) {
    if (h) h->err = msg;
    return s;
}

#define HIPCHK(h, call)                                                                  \
    do {                                                                                 \
        hipError_t e_ = (call);                                                          \
        if (e_ != hipSuccess)                                                            \
            return fail(h, CA_ERR_DEVICE, std::string(#call ": ") + hipGetErrorString(e_)); \
    } while (0)

template <class T>
ca_status dalloc(ca_handle *h, T **p, size_t count) {
    void *q = nullptr;
    hipError_t e = hipMalloc(&q, count * sizeof(T) + 16);
    if (e != hipSuccess) return fail(h, CA_ERR_DEVICE, std::string("hipMalloc: ") + hipGetErrorString(e));
    h->allocs.push_back(q);
    *p = (T *)q;
    return CA_OK;
}

// frees one of the handle's allocations (after the device is idle: kernels in
// flight on any stream may still read it) and drops it from the list
ca_status dfree(ca_handle *h, void *p) {
    if (!p) return CA_OK;
    HIPCHK(h, hipDeviceSynchronize());
    for (size_t i = 0; i < h->allocs.size(); ++i)
        if (h->allocs[i] == p) {
            h->allocs.erase(h->allocs.begin() + i);
            HIPCHK(h, hipFree(p));
            return CA_OK;
        }
    return fail(h, CA_ERR_ARG, "dfree: not an allocation of this handle");
}

constexpr uint32_t kNaNBits = 0xffffffffu;  // a NaN as float, and its pairs a NaN as double

ca_status fill_words(ca_handle *h, void *p, int64_t nwords, uint32_t v, hipStream_t st) {
    if (!p || nwords <= 0) return CA_OK;
    const int64_t blocks = std::min<int64_t>((nwords + 255) / 256, 4096);
    k_fill_u32<<<(unsigned)blocks, 256, 0, st>>>((uint32_t *)p, nwords, v);
    HIPCHK(h, hipGetLastError());
    return CA_OK;
}

// Debug poisoning: the forward work buffers (no entry point carries them from
// one call to the next) get NaN at every entry point while h->poison is set,
// so that a read of a value the call did not write shows up as NaN.  The
// Jacobian bands (update_jacobian -> ldiv) and the checkpoints and records
// (step_record -> step_pullback) are poisoned only by the calls that write them.
ca_status poison_work(ca_handle *h, hipStream_t st) {
    if (!h->poison) return CA_OK;
    const DevGeom &g = h->g;
    ca_status s;
    if ((s = fill_words(h, h->pre, 4 * g.nC, kNaNBits, st))) return s;
    if ((s = fill_words(h, h->L, 4 * g.nC, kNaNBits, st))) return s;
    if ((s = fill_words(h, h->scr, 2 * g.nF + 6 * g.nC, kNaNBits, st))) return s;
    for (int i = 0; i < 4; ++i)
        if ((s = fill_words(h, h->FE[i], h->S, kNaNBits, st))) return s;
    for (int i = 1; i < 4; ++i)
        if ((s = fill_words(h, h->FI[i], h->fi_size, kNaNBits, st))) return s;
    if ((s = fill_words(h, h->U, h->S, kNaNBits, st))) return s;
    return fill_words(h, h->raw, h->S, kNaNBits, st);
}

template <class T>
ca_status upload(ca_handle *h, T **dst, const std::vector<T> &src) {
    ca_status s = dalloc(h, dst, src.size() ? src.size() : 1);
    if (s) return s;
    if (!src.empty()) HIPCHK(h, hipMemcpy(*dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice));
    return CA_OK;
}

// ARS(3,4,3) (Ascher, Ruuth & Spiteri 1997) with the a42 = a43 explicit family
void ars343(ca_handle *h) {
    double g = 0.4358665215;
    for (int it = 0; it < 50; ++it) g -= (6 * g * g * g - 18 * g * g + 9 * g - 1) / (18 * g * g - 36 * g + 9);
    double b1 = -1.5 * g * g + 4 * g - 0.25, b2 = 1.5 * g * g - 5 * g + 1.25;
    double a42 = 0.5529291480359398, a43 = a42;
    double a31 = (1 - 4.5 * g + 1.5 * g * g) * a42 + (2.75 - 10.5 * g + 3.75 * g * g) * a43 - 3.5 + 13 * g - 4.5 * g * g;
    double a32 = (-1 + 4.5 * g - 1.5 * g * g) * a42 + (-2.75 + 10.5 * g - 3.75 * g * g) * a43 + 4 - 12.5 * g + 4.5 * g * g;
    memset(h->aE, 0, sizeof(h->aE));
    memset(h->aI, 0, sizeof(h->aI));
    h->aE[1][0] = g;
    h->aE[2][0] = a31;
    h->aE[2][1] = a32;
    h->aE[3][0] = 1 - a42 - a43;
    h->aE[3][1] = a42;
    h->aE[3][2] = a43;
    h->aI[1][1] = g;
    h->aI[2][1] = (1 - g) / 2;
    h->aI[2][2] = g;
    h->aI[3][1] = b1;
    h->aI[3][2] = b2;
    h->aI[3][3] = g;
    h->b[0] = 0;
    h->b[1] = b1;
    h->b[2] = b2;
    h->b[3] = g;
}

struct Launch {
    ca_handle *h;
    hipStream_t st;
    int id;
    hipEvent_t a = nullptr, b = nullptr;
    Launch(ca_handle *h_, hipStream_t st_, int id_) : h(h_), st(st_), id(id_) {
        if (h->prof) {
            (void)hipEventCreate(&a);
            (void)hipEventCreate(&b);
            (void)hipEventRecord(a, st);
        }
    }
    ~Launch() {
        if (h->prof) {
            (void)hipEventRecord(b, st);
            h->recs.push_back({id, a, b});
        }
    }
};

inline unsigned cdiv(int64_t a, int64_t b) { return (unsigned)((a + b - 1) / b); }

ca_status check_launch(ca_handle *h, const char *what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(h, CA_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
    return CA_OK;
}

// 1-D grid of (column block, level chunk) pairs for k_dss_gather
// k_dss_gather remaps block b to logical block (b % 8) * q + b / 8 with
// q = ceil(nb / 8): a bijection onto [0, 8 q) only if the grid holds 8 q
// blocks; logical blocks past nb return at once.
inline unsigned dss_grid(const DevGeom &g, int chunk) {
    const unsigned nb = cdiv(g.ncol, 256) * cdiv(g.nlev, chunk);
    return 8 * cdiv(nb, 8);
}

// copy src -> dst with the impenetrability filter (columns without DSS)
ca_status filter_columns(ca_handle *h, const float *src, float *dst, hipStream_t st) {
    const DevGeom &g = h->g;
    k_dss_gather<4, 1, true><<<dss_grid(g, g.nlev), 256, 0, st>>>(g, src, dst, 1, g.nlev);
    return check_launch(h, "filter");
}

// grid of the LDS-staged DSS: (cluster, level chunk) pairs, each of the 8
// XCDs taking ceil(ncl / 8) clusters for every level chunk
inline unsigned dss_lds_grid(const DevGeom &g, int chunk) {
    return 8 * cdiv(g.ncl, 8) * cdiv(g.nlev, chunk);
}

// DSS of NC centre planes (vector pair at VEC) and, with FACE, a face plane,
// src -> dst: the LDS-staged kernel on the sphere
// LDS-staged DSS; 16-byte staging loads when src is 16-byte aligned (the
// planes and rows of a sphere grid then are too: ncol is a multiple of 16)
template <int NC, int VEC, bool FACE>
void launch_dss_lds(const DevGeom &g, const float *src, float *dst, int filter, int ch, hipStream_t st) {
    if (((uintptr_t)src & 15) == 0)
        k_dss_lds<NC, VEC, FACE, true><<<dss_lds_grid(g, ch), 256, 0, st>>>(g, src, dst, filter, ch);
    else
        k_dss_lds<NC, VEC, FACE, false><<<dss_lds_grid(g, ch), 256, 0, st>>>(g, src, dst, filter, ch);
}

template <int NC, int VEC, bool FACE>
void launch_dss(const ca_handle *h, const float *src, float *dst, int filter, hipStream_t st) {
    const DevGeom &g = h->g;
    const int ch = h->dss_chunk > 0 ? h->dss_chunk : g.nlev;
    if (h->dss_kind == 0 && g.ncl > 0)
        launch_dss_lds<NC, VEC, FACE>(g, src, dst, filter, ch, st);
    else
        k_dss_gather<NC, VEC, FACE><<<dss_grid(g, ch), 256, 0, st>>>(g, src, dst, filter, ch);
}

// DSS of a full state-shaped buffer, src -> dst (out of place)
ca_status dss_gather_state(ca_handle *h, const float *src, float *dst, int filter, hipStream_t st, int kid) {
    {
        Launch l(h, st, kid);
        launch_dss<4, 1, true>(h, src, dst, filter, st);
    }
    return check_launch(h, "k_dss(state)");
}

// in-place state DSS (dss! hook): copy to the scratch buffer, gather back
ca_status dss_state(ca_handle *h, float *Y, hipStream_t st, int kid) {
    if (h->g.nelem == 0) return CA_OK;
    HIPCHK(h, hipMemcpyAsync(h->raw, Y, (size_t)h->S * sizeof(float), hipMemcpyDeviceToDevice, st));
    return dss_gather_state(h, h->raw, Y, 0, st, kid);
}

template <bool FIN>
void launch_explicit(ca_handle *h, const float *Y, const float *Lw, float *out, const FinFuse &fin, hipStream_t st) {
    const DevGeom &g = h->g;
    const DevNum &n = h->n;
    const int ch = FIN ? h->final_chunk : h->explicit_chunk;
    const int nchunks = ch > 0 ? (int)cdiv(g.nlev, ch) : 1;
    const dim3 grid(cdiv(g.ncol, 256), nchunks), blk(256);
    // with implicit diffusion the vertical diffusion is in the implicit tendency
    const bool hy = n.hyper != 0, vd = n.vdiff != 0 && !n.idiff;
    if (n.sphere) {
        if (hy && !vd) k_explicit<true, true, false, FIN><<<grid, blk, 0, st>>>(g, n, Y, Lw, out, ch, fin);
        else if (!hy && !vd) k_explicit<true, false, false, FIN><<<grid, blk, 0, st>>>(g, n, Y, Lw, out, ch, fin);
        else if (hy && vd) k_explicit<true, true, true, FIN><<<grid, blk, 0, st>>>(g, n, Y, Lw, out, ch, fin);
        else k_explicit<true, false, true, FIN><<<grid, blk, 0, st>>>(g, n, Y, Lw, out, ch, fin);
    } else {
        if (vd) k_explicit<false, false, true, FIN><<<grid, blk, 0, st>>>(g, n, Y, h->L, out, ch, fin);
        else k_explicit<false, false, false, FIN><<<grid, blk, 0, st>>>(g, n, Y, h->L, out, ch, fin);
    }
}

// remaining_tendency!: Laplacians, their DSS, the fused explicit kernel.
// With dss_out, the tendency is written un-DSSed to the scratch buffer and
// gathered (DSSed) into Yt; otherwise it goes to Yt as the reference's
// remaining_tendency! returns it.  With fin, Yt receives the step's final
// combination Y + fin's terms instead (un-DSSed; Yt must not be h->raw's
// Laplacian use, which ends before the explicit kernel starts).
ca_status explicit_tendency(ca_handle *h, float *Yt, const float *Y, hipStream_t st, bool dss_out = false,
                            float *Lout = nullptr, const FinFuse *fin = nullptr) {
    const DevGeom &g = h->g;
    const DevNum &n = h->n;
    const bool dss = n.sphere;
    float *Lw = Lout ? Lout : h->L;  // the DSSed Laplacians
    if (n.sphere && n.hyper) {
        {
            Launch l(h, st, KID_LAPLACIANS);
            k_laplacians<<<dim3(cdiv(g.ncol, 256), cdiv(g.nlev, h->lap_chunk)), 256, 0, st>>>(g, n, Y, h->raw,
                                                                                       h->lap_chunk);
        }
        {
            Launch l(h, st, KID_DSS_LAPLACIAN);
            launch_dss<4, 0, false>(h, h->raw, Lw, 0, st);
        }
    }
    if (fin) dss_out = false;
    float *out = (dss_out && dss) ? h->raw : Yt;
    {
        Launch l(h, st, fin ? KID_EXPLICIT_FINAL : KID_EXPLICIT);
        if (fin)
            launch_explicit<true>(h, Y, Lw, out, *fin, st);
        else
            launch_explicit<false>(h, Y, Lw, out, FinFuse{}, st);
        // surface fluxes at the first centre (surface_flux_tendency!, called
        // from additional_tendency!: part of the tendency the DSS averages)
        if (n.sflux) k_sfc_flux<<<cdiv(g.ncol, 256), 256, 0, st>>>(g, n, Y, out, fin ? fin->cg : 1.0f);
    }
    ca_status s = check_launch(h, "explicit tendency");
    if (s) return s;
    if (dss_out && dss) return dss_gather_state(h, h->raw, Yt, 0, st, KID_DSS_TENDENCY);
    return CA_OK;
}

inline bool ca_al16(const void *p) { return ((uintptr_t)p & 15) == 0; }

template <int TC, bool V4>
void launch_pcr_tcv(ca_handle *h, const StageSrc &src, float dtg, int i, float *U, float *Ut, hipStream_t st,
                    bool keep_fi) {
    float *FI = keep_fi ? h->FI[i] : nullptr;
    const DevGeom &g = h->g;
    const dim3 grid(cdiv(g.ncol, TC)), blk(256);
    if (src.nfull == 2 && src.nfi == 0)
        k_newton_pcr<TC, 2, 0, V4><<<grid, blk, 0, st>>>(g, h->n, src, dtg, U, FI, Ut);
    else if (src.nfull == 3 && src.nfi == 1)
        k_newton_pcr<TC, 3, 1, V4><<<grid, blk, 0, st>>>(g, h->n, src, dtg, U, FI, Ut);
    else if (src.nfull == 4 && src.nfi == 2)
        k_newton_pcr<TC, 4, 2, V4><<<grid, blk, 0, st>>>(g, h->n, src, dtg, U, FI, Ut);
    else
        k_newton_pcr<TC, -1, -1, V4><<<grid, blk, 0, st>>>(g, h->n, src, dtg, U, FI, Ut);
}

// 16-byte row accesses when every row of every plane is 16-byte aligned:
// ncol a multiple of 4 and all base pointers 16-byte aligned
template <int TC>
void launch_pcr_tc(ca_handle *h, const StageSrc &src, float dtg, int i, float *U, float *Ut, hipStream_t st,
                   bool keep_fi) {
    bool v4 = TC == 16 && h->g.ncol % 4 == 0 && ca_al16(U) && ca_al16(h->FI[i]) && (!Ut || ca_al16(Ut));
    for (int t = 0; t < src.nfull; ++t) v4 = v4 && ca_al16(src.full[t]);
    for (int t = 0; t < src.nfi; ++t) v4 = v4 && ca_al16(src.fi[t]);
    if (v4)
        launch_pcr_tcv<16, true>(h, src, dtg, i, U, Ut, st, keep_fi);
    else
        launch_pcr_tcv<TC, false>(h, src, dtg, i, U, Ut, st, keep_fi);
}

// the implicit stage i from src into U (h->U unless the caller records);
// Ut (optional) receives the assembled U~; keep_fi = false leaves F_I
// unstored (the step's last stage: the final combination does not read it)
void launch_newton_pcr(ca_handle *h, const StageSrc &src, float dtg, int i, hipStream_t st, float *U = nullptr,
                       float *Ut = nullptr, bool keep_fi = true) {
    if (!U) U = h->U;
    if (h->newton_variant == 2)
        launch_pcr_tc<32>(h, src, dtg, i, U, Ut, st, keep_fi);
    else
        launch_pcr_tc<16>(h, src, dtg, i, U, Ut, st, keep_fi);
}

ca_status ensure_bands(ca_handle *h) {
    if (h->bands) return CA_OK;
    const DevGeom &g = h->g;
    ca_status s = dalloc(h, &h->bands, (size_t)(4 * g.nC + 11 * g.nF));
    if (s) return s;
    float *p = h->bands;
    Bands &B = h->B;
    B.Aru0 = p; p += g.nC;
    B.Aru1 = p; p += g.nC;
    B.Aeu0 = p; p += g.nC;
    B.Aeu1 = p; p += g.nC;
    float **fp[] = {&B.Aur0, &B.Aur1, &B.Aue0, &B.Aue1, &B.Auh0a, &B.Auh0b, &B.Auh1a, &B.Auh1b, &B.lo, &B.d, &B.up};
    for (float **x : fp) {
        *x = p;
        p += g.nF;
    }
    return CA_OK;
}


// Per-level tables of one column grid, computed in double and stored in
// float (plus float64 copies of the geopotential differences): lc holds 8
// centre rows of nl (z, Jv, s, Rayleigh zeta, Phi, 1/s^2, 1/Jv, 1/(s^2 Jv)),
// lf 9 face rows of nl + 1 (z, Jv, s, dPhi, zeta, 1/Jv^2, 1/Jv, 1/s,
// 1/(s^2 Jv)).  R = 0: no deep-atmosphere scaling.
struct LevelTables {
    std::vector<float> lc, lf;
    std::vector<double> dphid, phicd;
};
LevelTables level_tables(int nl, const double *z_face, double R, double grav, double zd) {
    LevelTables T;
    std::vector<double> zf(z_face, z_face + nl + 1), zcd(nl);
    T.lc.assign(8 * (size_t)nl, 0.0f);
    T.lf.assign(9 * (size_t)(nl + 1), 0.0f);
    const double zmax = zf[nl];
    for (int k = 0; k < nl; ++k) zcd[k] = 0.5 * (zf[k] + zf[k + 1]);
    auto zeta = [&](double z) {
        if (!(z > zd)) return 0.0;
        const double sn = std::sin(M_PI * (z - zd) / (zmax - zd) / 2);
        return sn * sn;
    };
    for (int k = 0; k < nl; ++k) {
        const double sk = R > 0 ? (R + zcd[k]) / R : 1.0, Jv = zf[k + 1] - zf[k];
        const double v[8] = {zcd[k], Jv, sk, zeta(zcd[k]), grav * zcd[k], 1.0 / (sk * sk), 1.0 / Jv, 1.0 / (sk * sk * Jv)};
        for (int r = 0; r < 8; ++r) T.lc[r * (size_t)nl + k] = (float)v[r];
    }
    const int nf = nl + 1;
    for (int f = 0; f <= nl; ++f) {
        const double Jvf = f == 0 ? zf[1] - zf[0] : (f == nl ? zf[nl] - zf[nl - 1] : zcd[f] - zcd[f - 1]);
        const double sf = R > 0 ? (R + zf[f]) / R : 1.0;
        const double v[9] = {zf[f], Jvf, sf, f >= 1 && f < nl ? grav * (zcd[f] - zcd[f - 1]) : 0.0, zeta(zf[f]),
                             1.0 / (Jvf * Jvf), 1.0 / Jvf, 1.0 / sf, 1.0 / (sf * sf * Jvf)};
        for (int r = 0; r < 9; ++r) T.lf[r * (size_t)nf + f] = (float)v[r];
    }
    T.dphid.assign(nl + 1, 0.0);
    T.phicd.assign(nl, 0.0);
    for (int f = 1; f < nl; ++f) T.dphid[f] = grav * (zcd[f] - zcd[f - 1]);
    for (int k = 0; k < nl; ++k) T.phicd[k] = grav * zcd[k];
    return T;
}
// points g's per-level tables at device copies laid out as in LevelTables
void point_level_tables(DevGeom &g, int nl, const float *dlc, const float *dlf, const double *ddphi,
                        const double *dphic) {
    const int nf = nl + 1;
    g.dPhif_d = ddphi;
    g.Phic_d = dphic;
    g.zc = dlc;
    g.Jvc = dlc + nl;
    g.dzc = dlc + nl;
    g.sc = dlc + 2 * nl;
    g.zeta_c = dlc + 3 * nl;
    g.Phic = dlc + 4 * nl;
    g.is2c = dlc + 5 * nl;
    g.iJvc = dlc + 6 * nl;
    g.isJc = dlc + 7 * nl;
    g.zf = dlf;
    g.Jvf = dlf + nf;
    g.sf = dlf + 2 * nf;
    g.dPhif = dlf + 3 * nf;
    g.zeta_f = dlf + 4 * nf;
    g.g33f = dlf + 5 * nf;
    g.iJvf = dlf + 6 * nf;
    g.isf = dlf + 7 * nf;
    g.isJf = dlf + 8 * nf;
}

// ca_create's body: fills h; on failure the caller frees h
ca_status create_impl(const ca_geometry *geom, const ca_params *par, const ca_config *cfg, ca_handle *h) {
    const int nl = geom->nlev, nc = geom->ncol;
    HIPCHK(h, hipGetDevice(&h->device));
    DevGeom &g = h->g;
    g.nelem = geom->nelem;
    g.ncol = nc;
    g.nlev = nl;
    g.nC = (int64_t)nl * nc;
    g.nF = (int64_t)(nl + 1) * nc;
    h->S = 4 * g.nC + g.nF;
    auto tofloat = [](const double *p, size_t m) {
        std::vector<float> v(m);
        for (size_t i = 0; i < m; ++i) v[i] = (float)p[i];
        return v;
    };
    ca_status s;
    const float **cp[] = {&g.Jh, &g.gu11, &g.gu12, &g.gu22, &g.gd11, &g.gd12, &g.gd22, &g.f3n, &g.f12a, &g.f12b, &g.iJh};
    std::vector<float> colv(11 * (size_t)nc);
    for (int i = 0; i < nc; ++i) {
        colv[0 * (size_t)nc + i] = (float)geom->J_h[i];
        colv[10 * (size_t)nc + i] = (float)(1.0 / geom->J_h[i]);
        for (int t = 0; t < 3; ++t) {
            colv[(1 + t) * (size_t)nc + i] = (float)geom->gup[(size_t)t * nc + i];
            colv[(4 + t) * (size_t)nc + i] = (float)geom->gdn[(size_t)t * nc + i];
        }
        colv[7 * (size_t)nc + i] = (float)geom->f3_num[i];
        colv[8 * (size_t)nc + i] = (float)geom->f12[i];
        colv[9 * (size_t)nc + i] = (float)geom->f12[(size_t)nc + i];
    }
    float *dcol = nullptr;
    if ((s = upload(h, &dcol, colv))) return s;
    for (int t = 0; t < 11; ++t) *cp[t] = dcol + (size_t)t * nc;
    float *dbasis = nullptr;
    if ((s = upload(h, &dbasis, tofloat(geom->basis, 12 * (size_t)nc)))) return s;
    g.basis = dbasis;
    // vertical tables (computed in double, stored in float)
    {
        const LevelTables T = level_tables(nl, geom->z_face, geom->radius, par->grav, cfg->zd_rayleigh);
        double *ddphi = nullptr, *dphic = nullptr;
        float *dlc = nullptr, *dlf = nullptr;
        if ((s = upload(h, &ddphi, T.dphid))) return s;
        if ((s = upload(h, &dphic, T.phicd))) return s;
        if ((s = upload(h, &dlc, T.lc))) return s;
        if ((s = upload(h, &dlf, T.lf))) return s;
        point_level_tables(g, nl, dlc, dlf, ddphi, dphic);
    }
    g.aw_col = nullptr;
    g.D0_col = nullptr;
    g.Tsfc = nullptr;
    if (cfg->sfc_flux) {
        float *p = nullptr;
        if ((s = upload(h, &p, tofloat(cfg->T_sfc_col, nc)))) return s;
        g.Tsfc = p;
    }
    if (cfg->alpha_w_col) {
        float *p = nullptr;
        if ((s = upload(h, &p, tofloat(cfg->alpha_w_col, nc)))) return s;
        g.aw_col = p;
    }
    g.kdec = nullptr;
    if (cfg->vert_diff && cfg->D0_col) {
        float *p = nullptr;
        if ((s = upload(h, &p, tofloat(cfg->D0_col, nc)))) return s;
        g.D0_col = p;
    }
    if (cfg->vert_diff) {
        // DecayWithHeightDiffusion's exp(-(z_c - z_sfc) / H) per level (1 for H = inf)
        std::vector<float> dec(nl);
        for (int k = 0; k < nl; ++k) {
            const double zc = 0.5 * (geom->z_face[k] + geom->z_face[k + 1]);
            const double H = cfg->vert_diff_H;
            dec[k] = H > 0 && std::isfinite(H) ? (float)std::exp(-(zc - geom->z_face[0]) / H) : 1.0f;
        }
        float *p = nullptr;
        if ((s = upload(h, &p, dec))) return s;
        g.kdec = p;
    }
    // DSS tables (gather rows per column; element clusters on the sphere)
    {
        DssTable T;
        const std::string msg = ca_build_dss_table(geom->nelem, nc, geom->nelem > 0 ? geom->n_dss : 0, geom->dss_ptr,
                                                   geom->dss_col, geom->dss_w, geom->basis, T);
        if (!msg.empty()) return fail(h, CA_ERR_ARG, "ca_create: " + msg);
        int *d_tc = nullptr;
        float *d_tw = nullptr, *d_tM = nullptr;
        if ((s = upload(h, &d_tc, T.tc))) return s;
        if ((s = upload(h, &d_tw, T.tw))) return s;
        if ((s = upload(h, &d_tM, T.tM))) return s;
        g.gtc = reinterpret_cast<const int4 *>(d_tc);
        g.gtw = reinterpret_cast<const float4 *>(d_tw);
        g.gtM = reinterpret_cast<const float4 *>(d_tM);
        g.ncl = (int)T.nE.size();
        // the transposed (adjoint) DSS: same slots and clusters, own weights, transposed transforms
        DssTable TT;
        const std::string msgT = ca_build_dss_table(geom->nelem, nc, geom->nelem > 0 ? geom->n_dss : 0, geom->dss_ptr,
                                                    geom->dss_col, geom->dss_w, geom->basis, TT, true);
        if (!msgT.empty()) return fail(h, CA_ERR_ARG, "ca_create: " + msgT);
        float *dT_tw = nullptr, *dT_tM = nullptr, *dT_cw = nullptr, *dT_cM = nullptr;
        if ((s = upload(h, &dT_tw, TT.tw))) return s;
        if ((s = upload(h, &dT_tM, TT.tM))) return s;
        if (g.ncl > 0) {
            if ((s = upload(h, &dT_cw, TT.cw))) return s;
            if ((s = upload(h, &dT_cM, TT.cM))) return s;
        }
        if (g.ncl > 0) {
            int *d_list = nullptr, *d_nE = nullptr, *d_ccol = nullptr;
            uint16_t *d_slot = nullptr;
            float *d_cw = nullptr, *d_cM = nullptr;
            if ((s = upload(h, &d_list, T.list))) return s;
            if ((s = upload(h, &d_nE, T.nE))) return s;
            if ((s = upload(h, &d_ccol, T.ccol))) return s;
            if ((s = upload(h, &d_slot, T.slot))) return s;
            if ((s = upload(h, &d_cw, T.cw))) return s;
            if ((s = upload(h, &d_cM, T.cM))) return s;
            g.dl_list = d_list;
            g.dl_nE = d_nE;
            g.dl_col = d_ccol;
            g.dl_slot = reinterpret_cast<const ushort4 *>(d_slot);
            g.dl_w = reinterpret_cast<const float4 *>(d_cw);
            g.dl_M = reinterpret_cast<const float4 *>(d_cM);
        }
        h->gT = g;
        h->gT.gtw = reinterpret_cast<const float4 *>(dT_tw);
        h->gT.gtM = reinterpret_cast<const float4 *>(dT_tM);
        if (g.ncl > 0) {
            h->gT.dl_w = reinterpret_cast<const float4 *>(dT_cw);
            h->gT.dl_M = reinterpret_cast<const float4 *>(dT_cM);
        }
    }
    float W[16], iW[16], Dm[4][4];
    for (int j = 0; j < 4; ++j)
        for (int i = 0; i < 4; ++i) {
            W[4 * j + i] = (float)(geom->gll_w[i] * geom->gll_w[j]);
            iW[4 * j + i] = (float)(1.0 / (geom->gll_w[i] * geom->gll_w[j]));
        }
    for (int i = 0; i < 4; ++i)
        for (int k = 0; k < 4; ++k) Dm[i][k] = (float)geom->gll_D[4 * i + k];
    HIPCHK(h, hipMemcpyToSymbol(HIP_SYMBOL(c_W), W, sizeof(W)));
    HIPCHK(h, hipMemcpyToSymbol(HIP_SYMBOL(c_iW), iW, sizeof(iW)));
    HIPCHK(h, hipMemcpyToSymbol(HIP_SYMBOL(c_D), Dm, sizeof(Dm)));
    DevNum &n = h->n;
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
        return fail(h, CA_ERR_ARG, "ca_create: R_d / cp_d must be 2/7 to within 1e-6 (dry air; the float64 Exner "
                                   "function refines q^(2/7))");
    n.dTsr = par->T_surf_ref;
    n.dTmr = par->T_min_ref;
    n.dTmin = par->T_min_sgs;
    n.hyper = cfg->hyperdiff && geom->nelem > 0;
    n.nu_v = (float)cfg->nu4_vort;
    n.nu_s = (float)cfg->nu4_scal;
    n.dd = (float)cfg->div_damp;
    n.upw = cfg->upwinding;
    n.rayleigh = cfg->rayleigh;
    n.alpha_uh = (float)cfg->alpha_uh;
    n.alpha_w = (float)cfg->alpha_w;
    n.vdiff = cfg->vert_diff;
    n.sflux = cfg->sfc_flux != 0;
    n.idiff = cfg->implicit_diffusion && cfg->vert_diff;
    n.siters = cfg->solve_iters > 0 ? cfg->solve_iters : 1;
    n.vdk = cfg->vert_diff_model;
    h->vd_plain = !cfg->vert_diff ||
                  (cfg->vert_diff_model == 0 && !(cfg->vert_diff_H > 0 && std::isfinite(cfg->vert_diff_H)));
    n.CE = (float)cfg->vert_diff_C_E;
    n.za = (float)(0.5 * (geom->z_face[1] - geom->z_face[0]));
    n.Cd = (float)cfg->Cd;
    n.Ch = (float)cfg->Ch;
    n.gust = (float)cfg->gustiness;
    n.dt = (float)cfg->dt;
    n.sphere = geom->nelem > 0;
    // work buffers
    h->fi_size = n.idiff ? h->S : 2 * g.nC + g.nF;
    if (n.idiff) {
        float *eb = nullptr;
        if ((s = dalloc(h, &eb, 6 * (size_t)g.nC))) return s;
        h->E = DiffBands{eb, eb + g.nC, eb + 2 * g.nC, eb + 3 * g.nC, eb + 4 * g.nC, eb + 5 * g.nC};
        if ((s = dalloc(h, &h->idw, 9 * (size_t)g.nC + 7 * (size_t)g.nF))) return s;
    }
    if ((s = dalloc(h, &h->pre, 4 * (size_t)g.nC))) return s;
    if ((s = dalloc(h, &h->L, 4 * (size_t)g.nC))) return s;
    if ((s = dalloc(h, &h->scr, 2 * (size_t)g.nF + 6 * (size_t)g.nC))) return s;
    for (int i = 0; i < 4; ++i)
        if ((s = dalloc(h, &h->FE[i], (size_t)h->S))) return s;
    for (int i = 1; i < 4; ++i)
        if ((s = dalloc(h, &h->FI[i], (size_t)h->fi_size))) return s;
    if ((s = dalloc(h, &h->U, (size_t)h->S))) return s;
    if ((s = dalloc(h, &h->raw, (size_t)h->S))) return s;
    if ((s = dalloc(h, &h->flag, 1))) return s;
    HIPCHK(h, hipMemset(h->L, 0, 4 * (size_t)g.nC * sizeof(float)));
    // the work buffers start as NaN (all bits set): a read of a value no
    // kernel wrote comes out non-finite instead of as stale memory
    for (float *p : {h->pre, h->U, h->raw})
        HIPCHK(h, hipMemset(p, 0xff, (p == h->pre ? 4 * (size_t)g.nC : (size_t)h->S) * sizeof(float)));
    HIPCHK(h, hipMemset(h->scr, 0xff, (2 * (size_t)g.nF + 6 * (size_t)g.nC) * sizeof(float)));
    for (int i = 0; i < 4; ++i) HIPCHK(h, hipMemset(h->FE[i], 0xff, (size_t)h->S * sizeof(float)));
    for (int i = 1; i < 4; ++i)
        HIPCHK(h, hipMemset(h->FI[i], 0xff, (size_t)h->fi_size * sizeof(float)));
    ars343(h);
    h->newton_variant = (nl + 1 <= 64) ? 1 : 0;
    // level chunks sized so that small grids still fill the chip: about
    // 1024 wavefronts for the per-node sweeps, 1024 workgroups for the DSS.
    // The explicit tendency's cap of 7 levels: its wavefronts are long and
    // few (6 per SIMD fit), so the last partial round of them costs; at
    // h_elem 30, 63 levels, chunks of 7 (9 per column) fill the slots almost
    // exactly twice: 112.6 us against 114-115 us for chunks of 5 or 8.
    {
        const int64_t layer_waves = (nc + 63) / 64;
        const int64_t want = (1024 + layer_waves - 1) / layer_waves;
        h->explicit_chunk = (int)std::min<int64_t>(7, std::max<int64_t>(4, (nl + want - 1) / want));
        // the last explicit launch runs 4 wavefronts per SIMD; at h_elem 30,
        // 63 levels its chunks of 11 (6 per column) fill the slots about twice:
        // 165.8 us against 170.7 us with 7
        h->final_chunk = (int)std::min<int64_t>(11, std::max<int64_t>(4, (nl + want - 1) / want));
        h->adj_chunk = (int)std::min<int64_t>(8, std::max<int64_t>(4, (nl + want - 1) / want));
        h->lap_chunk = (int)std::min<int64_t>(4, std::max<int64_t>(2, (nl + want - 1) / want));
        if (g.ncl > 0) {
            const int64_t wantd = (1024 + g.ncl - 1) / g.ncl;
            // 6 levels per workgroup measured best at h_elem 30, 63 levels (the
            // staged DSS's three launches 2-3% faster than with 8)
            h->dss_chunk = (int)std::min<int64_t>(6, std::max<int64_t>(2, (nl + wantd - 1) / wantd));
        }
    }
    return CA_OK;
}

// The Nq = 4 Gauss-Lobatto-Legendre weights and derivative matrix: the
// kernels keep W, 1/W and D in __constant__ memory shared by every handle of
// the process, so ca_create accepts these tables only.
bool is_gll4(const double *w, const double *D) {
    const double x[4] = {-1.0, -1.0 / std::sqrt(5.0), 1.0 / std::sqrt(5.0), 1.0};
    const double wq[4] = {1.0 / 6.0, 5.0 / 6.0, 5.0 / 6.0, 1.0 / 6.0};
    double bw[4];
    for (int k = 0; k < 4; ++k) {
        double p = 1.0;
        for (int m = 0; m < 4; ++m)
            if (m != k) p *= x[k] - x[m];
        bw[k] = 1.0 / p;
    }
    for (int i = 0; i < 4; ++i) {
        if (std::fabs(w[i] - wq[i]) > 1e-12) return false;
        double diag = 0.0;
        for (int k = 0; k < 4; ++k) {
            if (k == i) continue;
            const double d = (bw[k] / bw[i]) / (x[i] - x[k]);
            diag -= d;
            if (std::fabs(D[4 * i + k] - d) > 1e-10) return false;
        }
        if (std::fabs(D[4 * i + i] - diag) > 1e-10) return false;
    }
    return true;
}

// ca_create's argument checks; an empty string when they pass
std::string check_create_args(const ca_geometry *geom, const ca_params *par, const ca_config *cfg) {
    if (!geom || !par || !cfg) return "null geometry, params or config";
    const int nl = geom->nlev, nc = geom->ncol;
    // kernels address a level row with 32-bit lane byte offsets: 4 * ncol < 2^31
    if (nl < 3 || nc < 1 || nc > (1 << 28)) return "need nlev >= 3 and 1 <= ncol <= 2^28";
    if (geom->nelem < 0 || (geom->nelem > 0 && nc != 16 * geom->nelem)) return "ncol must be 16 * nelem on a sphere";
    if (!geom->z_face || !geom->J_h || !geom->gup || !geom->gdn || !geom->f3_num || !geom->f12 || !geom->gll_w ||
        !geom->gll_D)
        return "null geometry table (z_face, J_h, gup, gdn, f3_num, f12, gll_w and gll_D are required)";
    if (geom->nelem > 0 && (!geom->basis || !geom->dss_ptr || geom->n_dss < 0 ||
                            (geom->n_dss > 0 && (!geom->dss_col || !geom->dss_w))))
        return "a sphere needs basis, dss_ptr, dss_col and dss_w";
    for (int k = 0; k < nl; ++k)
        if (!(geom->z_face[k + 1] > geom->z_face[k])) return "z_face must increase strictly";
    if (!is_gll4(geom->gll_w, geom->gll_D)) return "gll_w / gll_D must be the Nq = 4 GLL weights and derivative matrix";
    if (cfg->upwinding != 0 && cfg->upwinding != 1 && cfg->upwinding != 3)
        return "upwinding must be 0 (none), 1 (first_order) or 3 (vanleer_limiter)";
    if (cfg->vert_diff && cfg->vert_diff_model != 0 && cfg->vert_diff_model != 1)
        return "vert_diff_model must be 0 (DecayWithHeightDiffusion) or 1 (VerticalDiffusion)";
    if (cfg->vert_diff && cfg->vert_diff_model == 0 && !cfg->D0_col)
        return "DecayWithHeightDiffusion needs D0_col";
    if (cfg->vert_diff && cfg->vert_diff_model == 0 && !(cfg->vert_diff_H >= 0))
        return "DecayWithHeightDiffusion needs vert_diff_H > 0 (0 or INFINITY: no decay)";
    if (cfg->vert_diff && cfg->vert_diff_model == 1 && !(cfg->vert_diff_C_E >= 0 && std::isfinite(cfg->vert_diff_C_E)))
        return "VerticalDiffusion needs a finite vert_diff_C_E >= 0";
    if (cfg->sfc_flux && !cfg->T_sfc_col) return "surface fluxes need T_sfc_col";
    if (cfg->implicit_diffusion && !cfg->vert_diff) return "implicit diffusion needs vertical diffusion";
    if (cfg->implicit_diffusion && (cfg->solve_iters < 1 || cfg->solve_iters > 100))
        return "solve_iters must be in [1, 100]";
    if (cfg->sfc_flux && !(cfg->Cd >= 0 && cfg->Ch >= 0 && cfg->gustiness >= 0))
        return "surface fluxes need Cd, Ch and gustiness >= 0";
    return std::string();
}

// The implicit stage i with implicit vertical diffusion, composed from the
// hook kernels as ClimaTimeSteppers composes it from the hooks: U~ into
// h->raw, R = dtg T_imp(U~) (the Newton iteration starts at U~), the Jacobian
// at U~, dU = solve(A, R) into U, U = U~ - dU, the post-Newton upwind
// correction at U, and F_I = (U - U~) / dtg with all five components.
template <bool V4>
void launch_idiff_pcr(ca_handle *h, const StageSrc &src, float dtg, int i, float *U, hipStream_t st, bool keep_fi) {
    const DevGeom &g = h->g;
    const unsigned nb = cdiv(g.ncol, 16);
    float *FI = keep_fi ? h->FI[i] : nullptr;
    if (src.nfull == 2 && src.nfi == 0)
        k_newton_idiff_pcr<2, 0, V4><<<nb, 256, 0, st>>>(g, h->n, src, dtg, U, FI);
    else if (src.nfull == 3 && src.nfi == 1)
        k_newton_idiff_pcr<3, 1, V4><<<nb, 256, 0, st>>>(g, h->n, src, dtg, U, FI);
    else if (src.nfull == 4 && src.nfi == 2)
        k_newton_idiff_pcr<4, 2, V4><<<nb, 256, 0, st>>>(g, h->n, src, dtg, U, FI);
    else
        k_newton_idiff_pcr<-1, -1, V4><<<nb, 256, 0, st>>>(g, h->n, src, dtg, U, FI);
}

// keep_fi = false: the fused stage does not store F_I (the step's last stage,
// whose F_I nothing reads); the hook-kernel stage uses F_I as its scratch.
ca_status idiff_stage(ca_handle *h, const StageSrc &src, float dtg, int i, float *U, hipStream_t st,
                      bool keep_fi = true) {
    const DevGeom &g = h->g;
    const DevNum &n = h->n;
    ca_status s;
    if (h->newton_variant >= 1) {  // the fused wavefront-per-column stage (nlev + 1 <= 64)
        bool v4 = g.ncol % 4 == 0 && ca_al16(U) && ca_al16(h->FI[i]);
        for (int t = 0; t < src.nfull; ++t) v4 = v4 && ca_al16(src.full[t]);
        for (int t = 0; t < src.nfi; ++t) v4 = v4 && ca_al16(src.fi[t]);
        if (v4)
            launch_idiff_pcr<true>(h, src, dtg, i, U, st, keep_fi);
        else
            launch_idiff_pcr<false>(h, src, dtg, i, U, st, keep_fi);
        return check_launch(h, "implicit-diffusion stage");
    }
    if ((s = ensure_bands(h))) return s;
    float *Ut = h->raw, *T = h->FI[i];
    const unsigned nb = cdiv(h->S, 256);
    const dim3 cl(cdiv(g.ncol, 128), g.nlev), fl(cdiv(g.ncol, 128), g.nlev + 1);
    k_stage_sum<<<nb, 256, 0, st>>>(g, src, Ut);
    k_implicit_tendency<<<fl, 128, 0, st>>>(g, n, T, Ut);
    k_vdiff_add<<<cl, 128, 0, st>>>(g, n, Ut, T);
    k_update_jacobian<<<fl, 128, 0, st>>>(g, n, Ut, dtg, h->B);
    k_diffusion_jacobian<<<cl, 128, 0, st>>>(g, n, Ut, dtg, h->E);
    k_ldiv_approx<<<cdiv(g.ncol, 128), 128, 0, st>>>(g, h->B, h->E, n.siters, dtg, U, T, h->idw);
    k_newton_update<<<nb, 256, 0, st>>>(g, Ut, U, U);
    const float *C = nullptr;
    if (n.upw != 0) {
        k_correct<<<cl, 128, 0, st>>>(g, n, T, U);
        C = T;
    }
    k_stage_finish<<<nb, 256, 0, st>>>(g, Ut, C, dtg, U, h->FI[i]);
    return check_launch(h, "implicit-diffusion stage");
}

// Records of one step for the reverse pass: U_0 (the step's start state),
// U~_i and U_i of the three implicit stages, and the DSSed Laplacians L_i at
// U_0 .. U_3 (hyperdiffusion only).
inline int64_t rec_size(const ca_handle *h) { return 7 * h->S + 16 * h->g.nC; }
inline float *rec_U(const ca_handle *h, float *rec, int i) { return rec + (i == 0 ? 0 : 3 + i) * h->S; }
inline float *rec_Ut(const ca_handle *h, float *rec, int i) { return rec + (int64_t)i * h->S; }
inline float *rec_L(const ca_handle *h, float *rec, int i) { return rec + 7 * h->S + (int64_t)i * 4 * h->g.nC; }

// One ARS343 step of Y in place; Y is continuous and filtered on entry and on
// exit.  With rec, the step's records are written there as well.
ca_status one_step(ca_handle *h, float *Y, double dt, hipStream_t st, float *rec) {
    const DevGeom &g = h->g;
    ca_status s;
    // stage 1: explicit tendency at u (DSSed)
    // recording: U_0 copied, U_i and the Laplacians written into their record slots
    // The last stage's tendency is left un-DSSed: it enters only the final
    // combination, whose end-of-step DSS averages it there.  The weighted DSS
    // P is linear and P(P x) = P x, so P(c + b_4 dt F_E,4) equals
    // P(c + b_4 dt P F_E,4) up to rounding, and one DSS pass is saved.
    const int last = 3;
    // The final combination is fused into the last explicit tendency, which
    // writes u_{n+1} (before its DSS) to h->raw: ARS343's implicit tableau is
    // stiffly accurate (a_I[3][j] = b[j] for every j, set together in
    // ars343), so u + dt sum_j b_j (F_E,j + F_I,j) = U_4 + dt sum_j
    // (b_j - a_E[3][j]) F_E,j.
    FinFuse fin{};
    for (int j = 0; j < 3; ++j) {
        fin.fe[j] = h->FE[j];
        fin.c[j] = (float)(dt * (h->b[j] - h->aE[3][j]));
    }
    fin.cg = (float)(dt * h->b[3]);
    if (rec) HIPCHK(h, hipMemcpyAsync(rec_U(h, rec, 0), Y, (size_t)h->S * sizeof(float), hipMemcpyDeviceToDevice, st));
    if ((s = explicit_tendency(h, h->FE[0], Y, st, true, rec ? rec_L(h, rec, 0) : nullptr))) return s;
    for (int i = 1; i < 4; ++i) {
        StageSrc src{};
        src.full[0] = Y;
        src.cf[0] = 1.0f;
        src.nfull = 1;
        for (int j = 0; j < i; ++j)
            if (h->aE[i][j] != 0.0) {
                src.full[src.nfull] = h->FE[j];
                src.cf[src.nfull] = (float)(dt * h->aE[i][j]);
                ++src.nfull;
            }
        src.nfi = 0;
        for (int j = 1; j < i; ++j)
            if (h->aI[i][j] != 0.0) {
                src.fi[src.nfi] = h->FI[j];
                src.ci[src.nfi] = (float)(dt * h->aI[i][j]);
                ++src.nfi;
            }
        float dtg = (float)(dt * h->aI[i][i]);
        if (h->n.idiff) {  // the implicit stage from the hook kernels (no records: no reverse mode)
            {
                Launch l(h, st, KID_NEWTON_STAGE);
                if ((s = idiff_stage(h, src, dtg, i, h->U, st, i < last))) return s;
            }
            if (i < last) {
                if ((s = explicit_tendency(h, h->FE[i], h->U, st, true, nullptr))) return s;
            } else if ((s = explicit_tendency(h, h->raw, h->U, st, false, nullptr, &fin))) {
                return s;
            }
            continue;
        }
        const bool pcr = h->newton_variant == 1 || h->newton_variant == 2;
        if (rec && !pcr) {  // the PCR stage records U~ itself
            sadj::k_stage_state<<<dim3(cdiv(g.ncol, 256), g.nlev + 1), 256, 0, st>>>(g, src, rec_Ut(h, rec, i));
            if ((s = check_launch(h, "k_stage_state"))) return s;
        }
        float *Ui = rec ? rec_U(h, rec, i) : h->U;
        {
            Launch l(h, st, KID_NEWTON_STAGE);
            if (pcr)
                launch_newton_pcr(h, src, dtg, i, st, Ui, rec ? rec_Ut(h, rec, i) : nullptr, i < last);
            else
                k_newton_stage<<<cdiv(g.ncol, 128), 128, 0, st>>>(g, h->n, src, dtg, Ui, h->FI[i], h->scr);
        }
        if ((s = check_launch(h, "k_newton_stage"))) return s;
        if ((s = explicit_tendency(h, i < last ? h->FE[i] : h->raw, Ui, st, i < last,
                                   rec ? rec_L(h, rec, i) : nullptr, i < last ? nullptr : &fin)))
            return s;
    }
    // u_{n+1} (in h->raw), then dss! and the filter
    if (g.nelem > 0) {
        if ((s = dss_gather_state(h, h->raw, Y, 1, st, KID_DSS_STATE))) return s;
    } else {
        if ((s = filter_columns(h, h->raw, Y, st))) return s;
    }
    return CA_OK;
}

}  // namespace

extern "C" {

ca_status ca_create(const ca_geometry *geom, const ca_params *par, const ca_config *cfg, ca_handle **out) {
    if (!out) return CA_ERR_ARG;
    *out = nullptr;
    const std::string msg = check_create_args(geom, par, cfg);
    if (!msg.empty()) {
        g_create_error = "ca_create: " + msg;
        return CA_ERR_ARG;
    }
    ca_handle *h = new ca_handle();
    const ca_status s = create_impl(geom, par, cfg, h);
    if (s != CA_OK) {
        g_create_error = h->err.empty() ? "ca_create: invalid argument" : h->err;
        ca_destroy(h);
        return s;
    }
    *out = h;
    return CA_OK;
}

static void drop_step_graph(ca_handle *h);
ca_status ca_destroy(ca_handle *h) {
    if (!h) return CA_ERR_ARG;
    drop_step_graph(h);
    for (auto &r : h->recs) {
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
    }
    for (void *p : h->allocs) (void)hipFree(p);
    delete h;
    return CA_OK;
}

int64_t ca_state_size(const ca_handle *h) { return h ? h->S : 0; }

const char *ca_last_error(const ca_handle *h) { return h ? h->err.c_str() : g_create_error.c_str(); }

ca_status ca_set_precomputed(ca_handle *h, float *Y, void *stream) {
    if (!h || !Y) return CA_ERR_ARG;
    hipStream_t st = (hipStream_t)stream;
    const DevGeom &g = h->g;
    ca_status s;
    if ((s = poison_work(h, st))) return s;
    {
        Launch l(h, st, KID_HOOKS);
        k_precompute<<<dim3(cdiv(g.ncol, 128), g.nlev), 128, 0, st>>>(g, h->n, Y, h->pre, h->pre + g.nC,
                                                                       h->pre + 2 * g.nC, h->pre + 3 * g.nC);
    }
    return check_launch(h, "k_precompute");
}

ca_status ca_remaining_tendency(ca_handle *h, float *Yt, const float *Y, void *stream) {
    if (!h || !Yt || !Y) return CA_ERR_ARG;
    ca_status s;
    if ((s = poison_work(h, (hipStream_t)stream))) return s;
    return explicit_tendency(h, Yt, Y, (hipStream_t)stream);
}

ca_status ca_implicit_tendency(ca_handle *h, float *Yt, const float *Y, void *stream) {
    if (!h || !Yt || !Y) return CA_ERR_ARG;
    hipStream_t st = (hipStream_t)stream;
    const DevGeom &g = h->g;
    ca_status s;
    if ((s = poison_work(h, st))) return s;
    {
        Launch l(h, st, KID_HOOKS);
        k_implicit_tendency<<<dim3(cdiv(g.ncol, 128), g.nlev + 1), 128, 0, st>>>(g, h->n, Yt, Y);
        if (h->n.idiff) k_vdiff_add<<<dim3(cdiv(g.ncol, 128), g.nlev), 128, 0, st>>>(g, h->n, Y, Yt);
    }
    return check_launch(h, "k_implicit_tendency");
}

ca_status ca_update_jacobian(ca_handle *h, const float *Y, double dtgamma, void *stream) {
    if (!h || !Y) return CA_ERR_ARG;
    ca_status s = ensure_bands(h);
    if (s) return s;
    hipStream_t st = (hipStream_t)stream;
    const DevGeom &g = h->g;
    if ((s = poison_work(h, st))) return s;
    if (h->poison && (s = fill_words(h, h->bands, 4 * g.nC + 11 * g.nF, kNaNBits, st))) return s;
    {
        Launch l(h, st, KID_HOOKS);
        k_update_jacobian<<<dim3(cdiv(g.ncol, 128), g.nlev + 1), 128, 0, st>>>(g, h->n, Y, (float)dtgamma, h->B);
        if (h->n.idiff)
            k_diffusion_jacobian<<<dim3(cdiv(g.ncol, 128), g.nlev), 128, 0, st>>>(g, h->n, Y, (float)dtgamma, h->E);
    }
    return check_launch(h, "k_update_jacobian");
}

ca_status ca_ldiv(ca_handle *h, float *dY, const float *R, void *stream) {
    if (!h || !dY || !R) return CA_ERR_ARG;
    if (!h->bands) return fail(h, CA_ERR_ARG, "ca_ldiv called before ca_update_jacobian");
    hipStream_t st = (hipStream_t)stream;
    const DevGeom &g = h->g;
    ca_status s;
    if ((s = poison_work(h, st))) return s;
    {
        Launch l(h, st, KID_HOOKS);
        if (h->n.idiff && h->newton_variant >= 1)  // one wavefront per column (nlev + 1 <= 64)
            k_ldiv_approx_pcr<<<cdiv(g.ncol, 8), 256, 0, st>>>(g, h->B, h->E, h->n.siters, 1.0f, dY, R);
        else if (h->n.idiff)
            k_ldiv_approx<<<cdiv(g.ncol, 128), 128, 0, st>>>(g, h->B, h->E, h->n.siters, 1.0f, dY, R, h->idw);
        else
            k_ldiv<<<cdiv(g.ncol, 128), 128, 0, st>>>(g, h->B, dY, R, h->scr);
    }
    return check_launch(h, "k_ldiv");
}

ca_status ca_correct_implicit_advection(ca_handle *h, float *Yt, const float *Y, void *stream) {
    if (!h || !Yt || !Y) return CA_ERR_ARG;
    hipStream_t st = (hipStream_t)stream;
    const DevGeom &g = h->g;
    ca_status s;
    if ((s = poison_work(h, st))) return s;
    {
        Launch l(h, st, KID_HOOKS);
        k_correct<<<dim3(cdiv(g.ncol, 128), g.nlev), 128, 0, st>>>(g, h->n, Yt, Y);
    }
    return check_launch(h, "k_correct");
}

ca_status ca_dss(ca_handle *h, float *Y, void *stream) {
    if (!h || !Y) return CA_ERR_ARG;
    ca_status s;
    if ((s = poison_work(h, (hipStream_t)stream))) return s;
    return dss_state(h, Y, (hipStream_t)stream, KID_DSS_STATE);
}

// the step starts from a continuous, filtered state (the reference's state
// after dss! and cache! at the end of every step)
static ca_status step_prologue(ca_handle *h, float *Y, hipStream_t st) {
    const DevGeom &g = h->g;
    ca_status s;
    if (g.nelem > 0) {
        HIPCHK(h, hipMemcpyAsync(h->U, Y, (size_t)h->S * sizeof(float), hipMemcpyDeviceToDevice, st));
        if ((s = dss_gather_state(h, h->U, Y, 1, st, KID_DSS_STATE))) return s;
    } else {
        FinalSrc fs{};
        fs.dt = 0.0f;
        Launch l(h, st, KID_FINAL);
        k_combine<<<cdiv(cdiv(h->S, 4), 256), 256, 0, st>>>(g, Y, h->U, fs);
        if ((s = check_launch(h, "k_combine"))) return s;
        if ((s = filter_columns(h, h->U, Y, st))) return s;
    }
    return CA_OK;
}

static void drop_step_graph(ca_handle *h) {
    if (h->gexec) (void)hipGraphExecDestroy(h->gexec);
    h->gexec = nullptr;
    h->g_Y = nullptr;
}

ca_status ca_step(ca_handle *h, float *Y, double dt, int32_t nsteps, void *stream) {
    if (!h || !Y || nsteps < 0 || !(dt > 0)) return CA_ERR_ARG;
    hipStream_t st = (hipStream_t)stream;
    ca_status s;
    if ((s = poison_work(h, st))) return s;
    if ((s = step_prologue(h, Y, st))) return s;
    for (int stp = 0; stp < nsteps; ++stp)
        if ((s = one_step(h, Y, dt, st, nullptr))) return s;
    return CA_OK;
}

ca_status ca_step_graph(ca_handle *h, float *Y, double dt, int32_t nsteps, void *stream) {
    if (!h || !Y || nsteps < 0 || !(dt > 0)) return CA_ERR_ARG;
    if (!stream) return fail(h, CA_ERR_ARG, "ca_step_graph: needs a non-default stream (graph capture)");
    if (h->prof || h->poison)
        return fail(h, CA_ERR_ARG, "ca_step_graph: turn profiling and debug poisoning off (per-launch host work)");
    hipStream_t st = (hipStream_t)stream;
    ca_status s;
    if ((s = step_prologue(h, Y, st))) return s;
    if (nsteps == 0) return CA_OK;
    if (!h->gexec || h->g_Y != Y || h->g_dt != dt || h->g_st != st) {
        drop_step_graph(h);
        // buffers a stage allocates on first use are allocated before the capture
        if (h->n.idiff && (s = ensure_bands(h))) return s;
        HIPCHK(h, hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed));
        ca_status cs = one_step(h, Y, dt, st, nullptr);
        hipGraph_t graph = nullptr;
        const hipError_t e = hipStreamEndCapture(st, &graph);
        if (cs) {
            if (graph) (void)hipGraphDestroy(graph);
            return cs;
        }
        if (e != hipSuccess) return fail(h, CA_ERR_DEVICE, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
        const hipError_t ei = hipGraphInstantiate(&h->gexec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (ei != hipSuccess) {
            h->gexec = nullptr;
            return fail(h, CA_ERR_DEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
        }
        h->g_Y = Y;
        h->g_dt = dt;
        h->g_st = st;
    }
    for (int stp = 0; stp < nsteps; ++stp) HIPCHK(h, hipGraphLaunch(h->gexec, st));
    return CA_OK;
}

}  // extern "C"

// --------------------------------------------------------------- layout
__global__ void k_pack(DevGeom g, const float *src, ca_layout lay, int nlev_c, float *dst, int dir) {
    int col = blockIdx.x * blockDim.x + threadIdx.x;
    int k = blockIdx.y;
    if (col >= g.ncol || k >= nlev_c) return;
    int e, i, j;
    if (g.nelem > 0) {
        e = col >> 4;
        j = (col >> 2) & 3;
        i = col & 3;
    } else {
        e = col;
        i = j = 0;
    }
    int64_t so = lay.offset + k * lay.s_v + i * lay.s_i + j * lay.s_j + (int64_t)e * lay.s_h;
    int64_t d = (int64_t)k * g.ncol + col;
    if (dir == 0)
        dst[d] = src[so];
    else
        dst[so] = src[d];
}

__global__ void k_nonfinite(const float *Y, int64_t S, int *flag) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < S && !isfinite(Y[i])) atomicOr(flag, 1);
}

extern "C" {

ca_status ca_pack_state(ca_handle *h, const float *src, const ca_layout *lay, float *Y, void *stream) {
    if (!h || !src || !lay || !Y) return CA_ERR_ARG;
    hipStream_t st = (hipStream_t)stream;
    const DevGeom &g = h->g;
    for (int c = 0; c < 5; ++c) {
        int nl = c < 4 ? g.nlev : g.nlev + 1;
        k_pack<<<dim3(cdiv(g.ncol, 128), nl), 128, 0, st>>>(g, src, lay[c], nl, Y + (int64_t)c * g.nC, 0);
    }
    return check_launch(h, "k_pack");
}

ca_status ca_unpack_state(ca_handle *h, const float *Y, const ca_layout *lay, float *dst, void *stream) {
    if (!h || !dst || !lay || !Y) return CA_ERR_ARG;
    hipStream_t st = (hipStream_t)stream;
    const DevGeom &g = h->g;
    for (int c = 0; c < 5; ++c) {
        int nl = c < 4 ? g.nlev : g.nlev + 1;
        k_pack<<<dim3(cdiv(g.ncol, 128), nl), 128, 0, st>>>(g, Y + (int64_t)c * g.nC, lay[c], nl, dst, 1);
    }
    return check_launch(h, "k_unpack");
}

ca_status ca_check_finite(ca_handle *h, const float *Y, int32_t *flag, void *stream) {
    if (!h || !Y || !flag) return CA_ERR_ARG;
    hipStream_t st = (hipStream_t)stream;
    HIPCHK(h, hipMemsetAsync(h->flag, 0, sizeof(int), st));
    k_nonfinite<<<cdiv(h->S, 256), 256, 0, st>>>(Y, h->S, h->flag);
    ca_status s = check_launch(h, "k_nonfinite");
    if (s) return s;
    int v = 0;
    HIPCHK(h, hipMemcpyAsync(&v, h->flag, sizeof(int), hipMemcpyDeviceToHost, st));
    HIPCHK(h, hipStreamSynchronize(st));
    *flag = v;
    return v ? fail(h, CA_ERR_NONFINITE, "state holds NaN or Inf") : CA_OK;
}

ca_status ca_profile_enable(ca_handle *h, int32_t enable) {
    if (!h) return CA_ERR_ARG;
    HIPCHK(h, hipDeviceSynchronize());
    for (auto &r : h->recs) {
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
    }
    h->recs.clear();
    h->prof = enable != 0;
    return CA_OK;
}

int32_t ca_profile_count(void) { return KID_COUNT; }

ca_status ca_set_debug_poison(ca_handle *h, int32_t on) {
    if (!h) return CA_ERR_ARG;
    h->poison = on != 0;
    return CA_OK;
}

ca_status ca_set_tuning(ca_handle *h, int32_t knob, int32_t value) {
    if (!h) return CA_ERR_ARG;
    drop_step_graph(h);  // a captured step bakes in the launch schedules
    if (knob == CA_TUNE_NEWTON) {
        if (value < 0 || value > 2) return fail(h, CA_ERR_ARG, "CA_TUNE_NEWTON must be 0, 1 or 2");
        if (value > 0 && h->g.nlev + 1 > 64)
            return fail(h, CA_ERR_ARG, "wavefront-per-column Newton stage needs nlev + 1 <= 64");
        h->newton_variant = value;
        return CA_OK;
    }
    if (knob == CA_TUNE_EXPLICIT_CHUNK) {
        if (value < 0 || value > h->g.nlev) return fail(h, CA_ERR_ARG, "chunk must be in [0, nlev]");
        h->explicit_chunk = value;
        h->final_chunk = value;
        return CA_OK;
    }
    if (knob == CA_TUNE_DSS_CHUNK) {
        if (value < 1 || value > h->g.nlev) return fail(h, CA_ERR_ARG, "DSS chunk must be in [1, nlev]");
        h->dss_chunk = value;
        return CA_OK;
    }
    if (knob == CA_TUNE_DSS_KIND) {
        if (value != 0 && value != 1) return fail(h, CA_ERR_ARG, "DSS kind must be 0 (LDS-staged) or 1 (direct)");
        h->dss_kind = value;
        return CA_OK;
    }
    if (knob == CA_TUNE_LAP_CHUNK) {
        if (value < 1 || value > h->g.nlev) return fail(h, CA_ERR_ARG, "Laplacian chunk must be in [1, nlev]");
        h->lap_chunk = value;
        return CA_OK;
    }
    if (knob == CA_TUNE_ADJ_CHUNK) {
        if (value < 1 || value > h->g.nlev) return fail(h, CA_ERR_ARG, "adjoint chunk must be in [1, nlev]");
        if (h->aw_d0 && value < h->adj_chunk)
            return fail(h, CA_ERR_ARG, "adjoint chunk cannot shrink once the reverse-mode workspace exists");
        h->adj_chunk = value;
        return CA_OK;
    }
    return fail(h, CA_ERR_ARG, "unknown tuning knob");
}

ca_status ca_get_tuning(const ca_handle *h, int32_t knob, int32_t *value) {
    if (!h || !value) return CA_ERR_ARG;
    if (knob == CA_TUNE_NEWTON) *value = h->newton_variant;
    else if (knob == CA_TUNE_EXPLICIT_CHUNK) *value = h->explicit_chunk;
    else if (knob == CA_TUNE_DSS_CHUNK) *value = h->dss_chunk;
    else if (knob == CA_TUNE_DSS_KIND) *value = h->dss_kind;
    else if (knob == CA_TUNE_ADJ_CHUNK) *value = h->adj_chunk;
    else if (knob == CA_TUNE_LAP_CHUNK) *value = h->lap_chunk;
    else return CA_ERR_ARG;
    return CA_OK;
}

ca_status ca_profile_get(ca_handle *h, int32_t id, const char **name, int64_t *launches, double *total_ms) {
    if (!h || id < 0 || id >= KID_COUNT) return CA_ERR_ARG;
    HIPCHK(h, hipDeviceSynchronize());
    double tot = 0.0;
    int64_t cnt = 0;
    for (auto &r : h->recs)
        if (r.id == id) {
            float ms = 0.0f;
            HIPCHK(h, hipEventElapsedTime(&ms, r.a, r.b));
            tot += ms;
            ++cnt;
        }
    if (name) *name = kKernelNames[id];
    if (launches) *launches = cnt;
    if (total_ms) *total_ms = tot;
    return CA_OK;
}

}  // extern "C"
