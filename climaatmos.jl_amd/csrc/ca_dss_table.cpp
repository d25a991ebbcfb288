// ca_dss_table.cpp -- host construction of the weighted-DSS tables.
#include "ca_dss_table.h"

#include <algorithm>

namespace {

// Element clusters of the LDS-staged DSS.  Elements are adjacent if they
// share a DSS node.  A cluster stages its own elements and all their
// neighbours, so compact clusters read fewer halo rows.
//   * Panel blocks: when the elements are numbered panel by panel and row by
//     row on 6 panels of h x h (e = p h^2 + y h + x; checked against the edge
//     adjacency: e + 1 and e + h are edge neighbours inside a panel), the
//     clusters are the 4 x 4 blocks of each panel (the last row and column of
//     blocks narrower when 4 does not divide h).
//   * Otherwise clusters grow breadth-first over edge neighbours (elements
//     sharing a node of multiplicity 2) from the lowest unassigned element,
//     up to CA_DSS_OWN elements while own + neighbouring elements stay within
//     CA_DSS_EMAX.
// Every index the kernel uses is checked here.
std::string build_clusters(int ne, DssTable &T) {
    const int nc = 16 * ne;
    std::vector<std::vector<int>> adj(ne), eadj(ne);
    for (int c = 0; c < nc; ++c) {
        int m = 0;
        for (int t = 0; t < 4; ++t) m += T.tw[4 * (size_t)c + t] != 0.0f;
        for (int t = 0; t < 4; ++t) {
            if (T.tw[4 * (size_t)c + t] == 0.0f) continue;
            const int e2 = T.tc[4 * (size_t)c + t] / 16;
            if (e2 == c / 16) continue;
            adj[c / 16].push_back(e2);
            if (m == 2) eadj[c / 16].push_back(e2);
        }
    }
    for (int e = 0; e < ne; ++e) {
        for (auto *v : {&adj[e], &eadj[e]}) {
            std::sort(v->begin(), v->end());
            v->erase(std::unique(v->begin(), v->end()), v->end());
        }
        if ((int)adj[e].size() + 1 > CA_DSS_EMAX) return "an element has too many neighbours";
    }
    auto staged = [&](const std::vector<int> &own) {
        std::vector<int> u(own);
        for (int e : own) u.insert(u.end(), adj[e].begin(), adj[e].end());
        std::sort(u.begin(), u.end());
        u.erase(std::unique(u.begin(), u.end()), u.end());
        return u;
    };
    std::vector<int> owner(ne, -1), pos(ne, -1);
    // panel-block numbering check
    int hp = 0;
    while (6 * (hp + 1) * (hp + 1) <= ne) ++hp;
    bool panels = hp >= 1 && 6 * hp * hp == ne;
    auto is_edge = [&](int a, int b) { return std::binary_search(eadj[a].begin(), eadj[a].end(), b); };
    for (int e = 0; panels && e < ne; ++e) {
        const int x = e % hp, y = (e / hp) % hp;
        if ((x + 1 < hp && !is_edge(e, e + 1)) || (y + 1 < hp && !is_edge(e, e + hp))) panels = false;
    }
    std::vector<std::vector<int>> blocks;
    if (panels) {
        constexpr int B = 4;  // 4 x 4 = CA_DSS_OWN elements
        for (int p = 0; p < 6; ++p)
            for (int by = 0; by < hp; by += B)
                for (int bx = 0; bx < hp; bx += B) {
                    std::vector<int> own;
                    for (int y = by; y < std::min(by + B, hp); ++y)
                        for (int x = bx; x < std::min(bx + B, hp); ++x) own.push_back(p * hp * hp + y * hp + x);
                    if ((int)staged(own).size() > CA_DSS_EMAX) {
                        panels = false;
                        break;
                    }
                    blocks.push_back(own);
                }
        if (!panels) blocks.clear();
    }
    static_assert(CA_DSS_OWN == 16, "panel blocks are 4 x 4 elements");
    for (int e0 = 0, bi = 0; panels ? bi < (int)blocks.size() : e0 < ne; panels ? ++bi : ++e0) {
        if (!panels && owner[e0] >= 0) continue;
        const int cid = (int)T.nE.size();
        std::vector<int> own;
        if (panels) {
            own = blocks[bi];
            for (int e : own) owner[e] = cid;
        } else {
            own.push_back(e0);
            std::vector<int> queue(eadj[e0].begin(), eadj[e0].end());
            owner[e0] = cid;
            for (size_t qi = 0; qi < queue.size() && (int)own.size() < CA_DSS_OWN; ++qi) {
                const int cand = queue[qi];
                if (owner[cand] >= 0) continue;
                std::vector<int> trial(own);
                trial.push_back(cand);
                if ((int)staged(trial).size() > CA_DSS_EMAX) continue;
                own.push_back(cand);
                owner[cand] = cid;
                queue.insert(queue.end(), eadj[cand].begin(), eadj[cand].end());
            }
        }
        // staged list: own elements, then the other staged elements ascending
        std::vector<int> lst(own);
        for (int e : staged(own))
            if (owner[e] != cid) lst.push_back(e);
        if ((int)lst.size() > CA_DSS_EMAX) return "a cluster exceeds the staging size";
        for (size_t i = 0; i < lst.size(); ++i) pos[lst[i]] = (int)i;
        const size_t b = T.list.size();
        T.list.resize(b + CA_DSS_EMAX, 0);
        std::copy(lst.begin(), lst.end(), T.list.begin() + b);
        T.nE.push_back((int)lst.size());
        for (int t = 0; t < 256; ++t) {
            const int oi = t / 16, node = t % 16;
            if (oi >= (int)own.size()) {
                T.ccol.push_back(-1);
                T.slot.insert(T.slot.end(), 4, 0);
                T.cw.insert(T.cw.end(), 4, 0.0f);
                T.cM.insert(T.cM.end(), 16, 0.0f);
                continue;
            }
            const int c = 16 * own[oi] + node;
            T.ccol.push_back(c);
            for (int u = 0; u < 4; ++u) {
                const int sc = T.tc[4 * (size_t)c + u];
                const int p = pos[sc / 16];
                if (p < 0 || p >= (int)lst.size() || lst[p] != sc / 16) return "a copy lies outside its cluster's staged elements";
                T.slot.push_back((uint16_t)(16 * p + sc % 16));
                T.cw.push_back(T.tw[4 * (size_t)c + u]);
            }
            for (int u = 0; u < 16; ++u) T.cM.push_back(T.tM[16 * (size_t)c + u]);
        }
        for (int e : lst) pos[e] = -1;
    }
    std::vector<int> seen(nc, 0);
    for (int c : T.ccol)
        if (c >= 0) ++seen[c];
    for (int c = 0; c < nc; ++c)
        if (seen[c] != 1) return "cluster ownership is not a partition of the columns";
    return "";
}

}  // namespace

std::string ca_build_dss_table(int nelem, int ncol, int ngroups, const int32_t *dss_ptr, const int32_t *dss_col,
                               const double *dss_w, const double *basis, DssTable &T, bool transpose) {
    const int nc = ncol;
    std::vector<int> cgrp(nc, -1);
    for (int gi = 0; gi < ngroups; ++gi) {
        const int b0 = dss_ptr[gi], m = dss_ptr[gi + 1] - b0;
        if (m < 1 || m > 4) return "a DSS group has fewer than 1 or more than 4 members";
        for (int t = b0; t < b0 + m; ++t) {
            const int c = dss_col[t];
            if (c < 0 || c >= nc || cgrp[c] != -1) return "DSS groups do not partition distinct columns";
            cgrp[c] = gi;
        }
    }
    T.tc.assign(4 * (size_t)nc, 0);
    T.tw.assign(4 * (size_t)nc, 0.0f);
    T.tM.assign(16 * (size_t)nc, 0.0f);
    auto bas = [&](int plane, int c) { return basis[(size_t)plane * nc + c]; };
    for (int c = 0; c < nc; ++c) {
        for (int t = 0; t < 4; ++t) T.tc[4 * (size_t)c + t] = c;
        const int gi = cgrp[c];
        if (gi < 0) {
            T.tw[4 * (size_t)c] = 1.0f;
            T.tM[16 * (size_t)c + 0] = 1.0f;
            T.tM[16 * (size_t)c + 3] = 1.0f;
            continue;
        }
        const int b0 = dss_ptr[gi], m = dss_ptr[gi + 1] - b0;
        double wown = 0.0;
        for (int t = 0; t < m; ++t)
            if (dss_col[b0 + t] == c) wown = dss_w[b0 + t];
        for (int t = 0; t < m; ++t) {
            const int mcol = dss_col[b0 + t];
            // forward: y_c = sum_t w_t M(c <- c_t) x_{c_t}, M_ij = a_i(c) . a^j(c_t).
            // transpose: xbar_c = sum_t w_c M(c_t <- c)^T ybar_{c_t}, entry (j, i) = a_i(c_t) . a^j(c)
            const double w = transpose ? wown : dss_w[b0 + t];
            T.tc[4 * (size_t)c + t] = mcol;
            T.tw[4 * (size_t)c + t] = (float)w;
            double m00 = 0, m01 = 0, m10 = 0, m11 = 0;
            for (int d = 0; d < 3; ++d) {
                if (transpose) {
                    m00 += bas(6 + d, mcol) * bas(d, c);
                    m01 += bas(9 + d, mcol) * bas(d, c);
                    m10 += bas(6 + d, mcol) * bas(3 + d, c);
                    m11 += bas(9 + d, mcol) * bas(3 + d, c);
                } else {
                    m00 += bas(6 + d, c) * bas(d, mcol);
                    m01 += bas(6 + d, c) * bas(3 + d, mcol);
                    m10 += bas(9 + d, c) * bas(d, mcol);
                    m11 += bas(9 + d, c) * bas(3 + d, mcol);
                }
            }
            float *Mt = &T.tM[16 * (size_t)c + 4 * t];
            Mt[0] = (float)(w * m00);
            Mt[1] = (float)(w * m01);
            Mt[2] = (float)(w * m10);
            Mt[3] = (float)(w * m11);
        }
    }
    if (nelem > 0) {
        if (nc != 16 * nelem) return "ncol != 16 * nelem";
        return build_clusters(nelem, T);
    }
    return "";
}
