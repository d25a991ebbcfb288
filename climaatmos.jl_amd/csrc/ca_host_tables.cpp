// ca_host_tables.cpp -- C entry points over the host-side table builders,
// compiled with ca_dss_table.cpp into libcadycore_host.so (no HIP) so the
// CPU tests can check the tables the GPU kernels index with.
#include <cstring>

#include "ca_dss_table.h"

extern "C" {

void *ca_host_dss_build_t(int32_t nelem, int32_t ncol, int32_t ngroups, const int32_t *dss_ptr, const int32_t *dss_col,
                          const double *dss_w, const double *basis, int32_t transpose, char *err, int32_t errlen);

// Builds the DSS tables; returns an opaque pointer (NULL on failure, with
// the reason in err).
void *ca_host_dss_build(int32_t nelem, int32_t ncol, int32_t ngroups, const int32_t *dss_ptr, const int32_t *dss_col,
                        const double *dss_w, const double *basis, char *err, int32_t errlen) {
    return ca_host_dss_build_t(nelem, ncol, ngroups, dss_ptr, dss_col, dss_w, basis, 0, err, errlen);
}

// As ca_host_dss_build; transpose != 0 builds the adjoint DSS's tables.
void *ca_host_dss_build_t(int32_t nelem, int32_t ncol, int32_t ngroups, const int32_t *dss_ptr, const int32_t *dss_col,
                          const double *dss_w, const double *basis, int32_t transpose, char *err, int32_t errlen) {
    DssTable *T = new DssTable();
    const std::string msg =
        ca_build_dss_table(nelem, ncol, ngroups, dss_ptr, dss_col, dss_w, basis, *T, transpose != 0);
    if (!msg.empty()) {
        if (err && errlen > 0) {
            std::strncpy(err, msg.c_str(), (size_t)errlen - 1);
            err[errlen - 1] = 0;
        }
        delete T;
        return nullptr;
    }
    return T;
}

int32_t ca_host_dss_nclusters(const void *t) { return (int32_t)static_cast<const DssTable *>(t)->nE.size(); }

// Copies the tables out: tc/tw [4 ncol], tM [16 ncol]; list [ncl * EMAX],
// nE [ncl], ccol [ncl * 256], slot [4 ncl * 256], cw [4 ncl * 256],
// cM [16 ncl * 256] (cluster arrays may be NULL when there are none).
void ca_host_dss_get(const void *t, int32_t *tc, float *tw, float *tM, int32_t *list, int32_t *nE, int32_t *ccol,
                     uint16_t *slot, float *cw, float *cM) {
    const DssTable &T = *static_cast<const DssTable *>(t);
    auto cp = [](auto *dst, const auto &v) {
        if (dst && !v.empty()) std::memcpy(dst, v.data(), v.size() * sizeof(v[0]));
    };
    cp(tc, T.tc);
    cp(tw, T.tw);
    cp(tM, T.tM);
    cp(list, T.list);
    cp(nE, T.nE);
    cp(ccol, T.ccol);
    cp(slot, T.slot);
    cp(cw, T.cw);
    cp(cM, T.cM);
}

void ca_host_dss_free(void *t) { delete static_cast<DssTable *>(t); }

int32_t ca_host_dss_emax(void) { return CA_DSS_EMAX; }
}
