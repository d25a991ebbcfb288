// ca_hooks_fv.hip -- the ClimaTimeSteppers hooks on ClimaCore FieldVectors:
// every state argument is a ca_fieldvector (the parent arrays of Y.c and Y.f
// plus the strides of each component), so a caller hands over the fields it
// owns without converting them first.
//
//   ca_set_precomputed_quantities_fv          set_precomputed_quantities!(Y, p, t)   (cache!)
//                                             src/cache/precomputed_quantities.jl:928-931
//   ca_set_implicit_precomputed_quantities_fv set_implicit_precomputed_quantities!(Y, p, t)
//                                             (cache_imp!) precomputed_quantities.jl:698
//   ca_remaining_tendency_fv                  remaining_tendency!(Yt, Yt_lim, Y, p, t)
//                                             src/prognostic_equations/remaining_tendency.jl:48
//   ca_implicit_tendency_fv                   implicit_tendency!(Yt, Y, p, t)  implicit_tendency.jl:36
//   ca_update_jacobian_fv                     update_jacobian!(jac, Y, p, dtgamma, t)  jacobian.jl:74-75
//   ca_ldiv_fv                                ldiv!(dY, jac, R)  jacobian.jl:78-82
//   ca_correct_implicit_advection_fv          T_post_imp!(Yt, Y, p, t)  implicit_tendency.jl:322
//   ca_dss_fv                                 dss!(Y, p, t)  constrain_state.jl:59
//
// A FieldVector already in the native layout (level-major planes, column
// index 16 element + 4 j + i, contiguous components) is used in place.  Any
// other layout -- ClimaCore's VIJFH (level fastest) in particular -- goes
// through an LDS-tiled transpose into a native staging buffer of the handle
// and, for outputs, back: the horizontal kernels read 64 consecutive columns
// of one level per wavefront, which a level-fastest array cannot give them
// without that transpose.  Results are bitwise those of the native hooks.
#include "ca_kernels.h"

namespace {

// fields of a FieldVector in kernel arguments
struct FvArg {
    float *c, *f;
    ca_layout lay[5];
};

constexpr int kTL = 16;  // levels per transpose tile
constexpr int kTC = 64;  // columns per transpose tile

}  // namespace

// One (16 levels x 64 columns) tile of component p between a strided field
// and a native plane, through LDS.  dir 0: strided -> native, 1: native ->
// strided.  The strided side is walked levels fastest when its level stride
// is 1 (VIJFH), columns fastest otherwise; the native side columns fastest.
__global__ void __launch_bounds__(256) k_fv_tile(FvArg a, int p, int nl, int ncol, float *plane, int dir) {
    __shared__ float t[kTL][kTC + 1];
    const ca_layout L = a.lay[p];
    const float *base = p < 4 ? a.c : a.f;
    const int k0 = blockIdx.y * kTL, c0 = blockIdx.x * kTC, tid = threadIdx.x;
    const bool lev_fast = L.s_v == 1;
    auto src_index = [&](int lv, int cc) {
        const int col = c0 + cc, k = k0 + lv;
        const int h = col >> 4, j = (col >> 2) & 3, i = col & 3;
        return L.offset + (int64_t)k * L.s_v + (int64_t)i * L.s_i + (int64_t)j * L.s_j + (int64_t)h * L.s_h;
    };
    if (dir == 0) {
        for (int q = tid; q < kTL * kTC; q += 256) {
            const int lv = lev_fast ? q % kTL : q / kTC, cc = lev_fast ? q / kTL : q % kTC;
            if (k0 + lv < nl && c0 + cc < ncol) t[lv][cc] = base[src_index(lv, cc)];
        }
        __syncthreads();
        for (int q = tid; q < kTL * kTC; q += 256) {
            const int lv = q / kTC, cc = q % kTC;
            if (k0 + lv < nl && c0 + cc < ncol) plane[(int64_t)(k0 + lv) * ncol + c0 + cc] = t[lv][cc];
        }
    } else {
        for (int q = tid; q < kTL * kTC; q += 256) {
            const int lv = q / kTC, cc = q % kTC;
            if (k0 + lv < nl && c0 + cc < ncol) t[lv][cc] = plane[(int64_t)(k0 + lv) * ncol + c0 + cc];
        }
        __syncthreads();
        float *out = p < 4 ? a.c : a.f;
        for (int q = tid; q < kTL * kTC; q += 256) {
            const int lv = lev_fast ? q % kTL : q / kTC, cc = lev_fast ? q / kTL : q % kTC;
            if (k0 + lv < nl && c0 + cc < ncol) out[src_index(lv, cc)] = t[lv][cc];
        }
    }
}

// zero of one strided component (remaining_tendency!'s Yt_lim .= 0)
__global__ void __launch_bounds__(256) k_fv_zero(FvArg a, int p, int nl, int ncol) {
    const int col = blockIdx.x * blockDim.x + threadIdx.x, k = blockIdx.y;
    if (col >= ncol || k >= nl) return;
    const ca_layout L = a.lay[p];
    float *base = p < 4 ? a.c : a.f;
    const int h = col >> 4, j = (col >> 2) & 3, i = col & 3;
    base[L.offset + (int64_t)k * L.s_v + (int64_t)i * L.s_i + (int64_t)j * L.s_j + (int64_t)h * L.s_h] = 0.0f;
}

namespace {

FvArg fv_arg(const ca_fieldvector *v) {
    FvArg a;
    a.c = v->c;
    a.f = v->f;
    for (int p = 0; p < 5; ++p) a.lay[p] = v->lay[p];
    return a;
}

// the native buffer a FieldVector is, or null
float *fv_native(const ca_handle *h, const ca_fieldvector *v) {
    const DevGeom &g = h->g;
    if (!v->c || v->f != v->c + 4 * g.nC) return nullptr;
    for (int p = 0; p < 5; ++p) {
        const ca_layout &L = v->lay[p];
        if (L.offset != (p < 4 ? p * g.nC : 0) || L.s_v != g.ncol || L.s_i != 1 || L.s_j != 4 || L.s_h != 16)
            return nullptr;
    }
    return v->c;
}

// argument checks: non-null parent arrays, non-negative offsets and strides
// (the extent of the caller's arrays is the caller's to guarantee)
ca_status fv_check(ca_handle *h, const ca_fieldvector *v, const char *what) {
    if (!v || !v->c || !v->f) return fail(h, CA_ERR_ARG, std::string(what) + ": null FieldVector or parent array");
    for (int p = 0; p < 5; ++p) {
        const ca_layout &L = v->lay[p];
        if (L.offset < 0 || L.s_v < 0 || L.s_i < 0 || L.s_j < 0 || L.s_h < 0)
            return fail(h, CA_ERR_ARG, std::string(what) + ": negative offset or stride");
    }
    return CA_OK;
}

ca_status fv_staging(ca_handle *h) {
    ca_status s;
    if (!h->fv_a && (s = dalloc(h, &h->fv_a, (size_t)h->S))) return s;
    if (!h->fv_b && (s = dalloc(h, &h->fv_b, (size_t)h->S))) return s;
    return CA_OK;
}

// strided -> native (dir 0) or native -> strided (dir 1), components [p0, p1)
ca_status fv_move(ca_handle *h, const ca_fieldvector *v, float *Y, int dir, hipStream_t st, int p0 = 0, int p1 = 5) {
    const DevGeom &g = h->g;
    const FvArg a = fv_arg(v);
    for (int p = p0; p < p1; ++p) {
        const int nl = p < 4 ? g.nlev : g.nlev + 1;
        k_fv_tile<<<dim3(cdiv(g.ncol, kTC), cdiv(nl, kTL)), 256, 0, st>>>(a, p, nl, g.ncol, Y + (int64_t)p * g.nC,
                                                                          dir);
    }
    return check_launch(h, dir == 0 ? "FieldVector gather" : "FieldVector scatter");
}

// the native buffer holding the input v: v itself, or staging buffer `stage`
// filled from it
ca_status fv_in(ca_handle *h, const ca_fieldvector *v, float *stage, hipStream_t st, float **out) {
    if ((*out = fv_native(h, v))) return CA_OK;
    ca_status s;
    if ((s = fv_staging(h))) return s;
    *out = stage == nullptr ? h->fv_a : stage;
    return fv_move(h, v, *out, 0, st);
}

}  // namespace

extern "C" {

ca_status ca_fieldvector_native(const ca_handle *h, float *Y, ca_fieldvector *out) {
    if (!h || !Y || !out) return CA_ERR_ARG;
    const DevGeom &g = h->g;
    out->c = Y;
    out->f = Y + 4 * g.nC;
    for (int p = 0; p < 5; ++p) out->lay[p] = ca_layout{p < 4 ? p * g.nC : 0, g.ncol, 1, 4, 16};
    return CA_OK;
}

ca_status ca_fieldvector_vijfh(const ca_handle *h, float *c, float *f, ca_fieldvector *out) {
    if (!h || !c || !f || !out) return CA_ERR_ARG;
    const int64_t nv = h->g.nlev, nf = nv + 1;
    out->c = c;
    out->f = f;
    // Julia column-major (Nv, Ni, Nj, F, Nh): v fastest, then i, j, component, element
    for (int p = 0; p < 4; ++p) out->lay[p] = ca_layout{p * nv * 16, 1, nv, 4 * nv, 4 * 16 * nv};
    out->lay[4] = ca_layout{0, 1, nf, 4 * nf, 16 * nf};
    return CA_OK;
}

ca_status ca_set_implicit_precomputed_quantities_fv(ca_handle *h, const ca_fieldvector *Y, double t, void *stream) {
    (void)t;
    if (!h) return CA_ERR_ARG;
    ca_status s;
    if ((s = fv_check(h, Y, "ca_set_implicit_precomputed_quantities_fv"))) return s;
    hipStream_t st = (hipStream_t)stream;
    if (float *y = fv_native(h, Y)) return ca_set_precomputed(h, y, stream);
    float *y = nullptr;
    if ((s = fv_in(h, Y, nullptr, st, &y))) return s;
    if ((s = ca_set_precomputed(h, y, stream))) return s;
    // the filter changes u_3 only
    return fv_move(h, Y, y, 1, st, 4, 5);
}

ca_status ca_set_precomputed_quantities_fv(ca_handle *h, const ca_fieldvector *Y, double t, void *stream) {
    // the explicit precomputed quantities are formed inside the kernels that
    // use them; what the caller sees of both calls is the filtered Y
    return ca_set_implicit_precomputed_quantities_fv(h, Y, t, stream);
}

ca_status ca_remaining_tendency_fv(ca_handle *h, const ca_fieldvector *Yt, const ca_fieldvector *Yt_lim,
                                   const ca_fieldvector *Y, double t, void *stream) {
    (void)t;
    if (!h) return CA_ERR_ARG;
    ca_status s;
    if ((s = fv_check(h, Yt, "ca_remaining_tendency_fv (Yt)")) || (s = fv_check(h, Y, "ca_remaining_tendency_fv (Y)")))
        return s;
    if (Yt_lim && (s = fv_check(h, Yt_lim, "ca_remaining_tendency_fv (Yt_lim)"))) return s;
    hipStream_t st = (hipStream_t)stream;
    const DevGeom &g = h->g;
    // Yt_lim .= 0: the dry model has no limited (tracer) tendencies
    if (Yt_lim) {
        const FvArg a = fv_arg(Yt_lim);
        for (int p = 0; p < 5; ++p) {
            const int nl = p < 4 ? g.nlev : g.nlev + 1;
            k_fv_zero<<<dim3(cdiv(g.ncol, 256), nl), 256, 0, st>>>(a, p, nl, g.ncol);
        }
        if ((s = check_launch(h, "Yt_lim zero"))) return s;
    }
    float *y = nullptr;
    if ((s = fv_in(h, Y, nullptr, st, &y))) return s;
    float *yt = fv_native(h, Yt);
    if (yt) return ca_remaining_tendency(h, yt, y, stream);
    if ((s = fv_staging(h))) return s;
    if ((s = ca_remaining_tendency(h, h->fv_b, y, stream))) return s;
    return fv_move(h, Yt, h->fv_b, 1, st);
}

ca_status ca_implicit_tendency_fv(ca_handle *h, const ca_fieldvector *Yt, const ca_fieldvector *Y, double t,
                                  void *stream) {
    (void)t;
    if (!h) return CA_ERR_ARG;
    ca_status s;
    if ((s = fv_check(h, Yt, "ca_implicit_tendency_fv (Yt)")) || (s = fv_check(h, Y, "ca_implicit_tendency_fv (Y)")))
        return s;
    hipStream_t st = (hipStream_t)stream;
    float *y = nullptr;
    if ((s = fv_in(h, Y, nullptr, st, &y))) return s;
    if (float *yt = fv_native(h, Yt)) return ca_implicit_tendency(h, yt, y, stream);
    if ((s = fv_staging(h))) return s;
    if ((s = ca_implicit_tendency(h, h->fv_b, y, stream))) return s;
    return fv_move(h, Yt, h->fv_b, 1, st);
}

ca_status ca_update_jacobian_fv(ca_handle *h, const ca_fieldvector *Y, double dtgamma, double t, void *stream) {
    (void)t;
    if (!h) return CA_ERR_ARG;
    ca_status s;
    if ((s = fv_check(h, Y, "ca_update_jacobian_fv"))) return s;
    float *y = nullptr;
    if ((s = fv_in(h, Y, nullptr, (hipStream_t)stream, &y))) return s;
    return ca_update_jacobian(h, y, dtgamma, stream);
}

ca_status ca_ldiv_fv(ca_handle *h, const ca_fieldvector *dY, const ca_fieldvector *R, void *stream) {
    if (!h) return CA_ERR_ARG;
    ca_status s;
    if ((s = fv_check(h, dY, "ca_ldiv_fv (dY)")) || (s = fv_check(h, R, "ca_ldiv_fv (R)"))) return s;
    hipStream_t st = (hipStream_t)stream;
    float *r = nullptr;
    if ((s = fv_in(h, R, nullptr, st, &r))) return s;
    if (float *dy = fv_native(h, dY)) return ca_ldiv(h, dy, r, stream);
    if ((s = fv_staging(h))) return s;
    if ((s = ca_ldiv(h, h->fv_b, r, stream))) return s;
    return fv_move(h, dY, h->fv_b, 1, st);
}

ca_status ca_correct_implicit_advection_fv(ca_handle *h, const ca_fieldvector *Yt, const ca_fieldvector *Y, double t,
                                           void *stream) {
    (void)t;
    if (!h) return CA_ERR_ARG;
    ca_status s;
    if ((s = fv_check(h, Yt, "ca_correct_implicit_advection_fv (Yt)")) ||
        (s = fv_check(h, Y, "ca_correct_implicit_advection_fv (Y)")))
        return s;
    hipStream_t st = (hipStream_t)stream;
    float *y = nullptr;
    if ((s = fv_in(h, Y, nullptr, st, &y))) return s;
    if (float *yt = fv_native(h, Yt)) return ca_correct_implicit_advection(h, yt, y, stream);
    if ((s = fv_staging(h))) return s;
    if ((s = ca_correct_implicit_advection(h, h->fv_b, y, stream))) return s;
    return fv_move(h, Yt, h->fv_b, 1, st);
}

ca_status ca_dss_fv(ca_handle *h, const ca_fieldvector *Y, double t, void *stream) {
    (void)t;
    if (!h) return CA_ERR_ARG;
    ca_status s;
    if ((s = fv_check(h, Y, "ca_dss_fv"))) return s;
    hipStream_t st = (hipStream_t)stream;
    if (float *y = fv_native(h, Y)) return ca_dss(h, y, stream);
    float *y = nullptr;
    if ((s = fv_in(h, Y, nullptr, st, &y))) return s;
    if ((s = ca_dss(h, y, stream))) return s;
    return fv_move(h, Y, y, 1, st);
}

}  // extern "C"
