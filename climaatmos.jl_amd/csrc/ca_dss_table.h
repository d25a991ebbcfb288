// ca_dss_table.h -- host construction of the weighted-DSS tables (plain C++,
// no HIP): the per-column gather rows and the element clusters of the
// LDS-staged kernel.  Built into libcadycore and, on its own, into a host
// library the CPU tests load.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#define CA_DSS_OWN 16   // own elements per DSS cluster (256 columns, one workgroup)
#define CA_DSS_EMAX 48  // staged elements per cluster, own + neighbours

struct DssTable {
    // per column, 4 slots: the copies of the column's node in the group's
    // canonical order, padded with (own column, weight 0); an interior node
    // is its own slot of weight 1.  tM: per slot w (a_i(own) . a^j(copy)),
    // i, j = 1, 2 row-major.
    std::vector<int> tc;
    std::vector<float> tw, tM;
    // clusters (sphere only): list[cl * EMAX + i] staged elements (own first),
    // nE[cl] their count; per (cl, thread): ccol (-1: none), slot (4 indices
    // list position * 16 + node), cw (4 weights), cM (16 transform entries)
    std::vector<int> list, nE, ccol;
    std::vector<uint16_t> slot;
    std::vector<float> cw, cM;
};

// Returns "" on success, otherwise what is wrong with the input.
// dss_ptr / dss_col / dss_w: shared-node groups (CSR, member columns,
// weights WJ / sum WJ); basis: 12 planes of ncol (a^1, a^2, a_1, a_2 as
// Cartesian x, y, z); nelem = 0 for independent columns.  transpose: the
// tables of the transposed (adjoint) DSS instead -- the same slots, every
// slot weighted by the own column's weight, each transform the transpose of
// the one the slot's column applies to the own column's copy.  The
// clusters and slot positions are the same as the forward table's.
std::string ca_build_dss_table(int nelem, int ncol, int ngroups, const int32_t *dss_ptr, const int32_t *dss_col,
                               const double *dss_w, const double *basis, DssTable &out, bool transpose = false);
