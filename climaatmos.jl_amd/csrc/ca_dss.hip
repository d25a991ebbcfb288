// ca_dss.hip -- weighted direct stiffness summation (ClimaCore
// Spaces.weighted_dss!, called at src/prognostic_equations/constrain_state.jl:59-64
// and remaining_tendency.jl:18-21).  Each shared node's value becomes the
// J-weighted average over its copies; covariant horizontal vectors are
// averaged as Cartesian vectors and projected back onto each element's basis.
// Thread per (shared-node group, level); interior nodes are untouched.
#include "ca_kernels.h"

__global__ void __launch_bounds__(256) k_dss_scalar(DevGeom g, float *f, int nplanes, int nlev_plane,
                                                    int64_t plane_stride) {
    int grp = blockIdx.x * blockDim.x + threadIdx.x;
    int k = blockIdx.y;
    if (grp >= g.ndss || k >= nlev_plane) return;
    const int b = g.dptr[grp], e = g.dptr[grp + 1];
    const int64_t off = (int64_t)k * g.ncol;
    for (int pl = 0; pl < nplanes; ++pl) {
        float *fk = f + pl * plane_stride + off;
        float s = 0.0f;
        for (int t = b; t < e; ++t) s = fmaf(g.dw[t], fk[g.dcol[t]], s);
        for (int t = b; t < e; ++t) fk[g.dcol[t]] = s;
    }
}

__global__ void __launch_bounds__(256) k_dss_vector(DevGeom g, float *u1, float *u2, int nlev_plane) {
    int grp = blockIdx.x * blockDim.x + threadIdx.x;
    int k = blockIdx.y;
    if (grp >= g.ndss || k >= nlev_plane) return;
    const int b = g.dptr[grp], e = g.dptr[grp + 1];
    const int64_t off = (int64_t)k * g.ncol;
    const int64_t nc = g.ncol;
    const float *B = g.basis;
    float X0 = 0.0f, X1 = 0.0f, X2 = 0.0f;
    for (int t = b; t < e; ++t) {
        int c = g.dcol[t];
        float w = g.dw[t];
        float a = w * u1[off + c], bb = w * u2[off + c];
        X0 += a * B[0 * nc + c] + bb * B[3 * nc + c];
        X1 += a * B[1 * nc + c] + bb * B[4 * nc + c];
        X2 += a * B[2 * nc + c] + bb * B[5 * nc + c];
    }
    for (int t = b; t < e; ++t) {
        int c = g.dcol[t];
        u1[off + c] = X0 * B[6 * nc + c] + X1 * B[7 * nc + c] + X2 * B[8 * nc + c];
        u2[off + c] = X0 * B[9 * nc + c] + X1 * B[10 * nc + c] + X2 * B[11 * nc + c];
    }
}

// Gather form of the weighted DSS, out of place: thread per (column, chunk
// of levels).  A column of a shared node reads every copy of its node and
// writes only itself; a column of an interior node copies.  Loads are
// issued own-column first (coalesced across the wavefront); the scalar sum
// runs over the copies in the group's canonical order (the rank of each
// slot), so every copy of a node computes bitwise the same value.
// Covariant horizontal vector planes (vec, vec + 1) use the per-slot 2x2
// transforms w_s (a_i(own) . a^j(slot)), i.e. the Cartesian average
// projected back on the column's own basis.  The per-column table is read
// once per thread (coalesced slot-major planes) and reused over its levels.
//   ncp: centre planes; vec: index of the first of the two vector planes
//   (-1: none); face: 1 if a face plane follows the centre planes;
//   filter: zero the bottom / top face values (impenetrability).
//   The 1-D grid covers (column block, level chunk) pairs; blocks are
//   dispatched round-robin over the 8 XCDs, so the logical order is remapped
//   to give every XCD a contiguous span of column blocks (neighbouring
//   elements, whose copies a gather reads, then share that XCD's L2).
__device__ __forceinline__ bool xcd_block(int nb_col, int nb_lev, int &cb, int &lb) {
    const int nb = nb_col * nb_lev;
    const int q = (nb + 7) / 8;
    const int b = blockIdx.x;
    const int logical = (b & 7) * q + (b >> 3);
    if (logical >= nb) return false;
    cb = logical / nb_lev;
    lb = logical - cb * nb_lev;
    return true;
}

// canonical-order weighted sum of up to four slot values
__device__ __forceinline__ float ranked_sum(const float v[4], const float w[4], int rank, int m) {
    float s = 0.0f;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        if (r >= m) break;
        // slot holding canonical rank r
        float vv = 0.0f, ww = 0.0f;
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const bool hit = ((rank >> (2 * t)) & 3) == r && t < m;
            vv = hit ? v[t] : vv;
            ww = hit ? w[t] : ww;
        }
        s = fmaf(ww, vv, s);
    }
    return s;
}

__global__ void __launch_bounds__(256) k_dss_gather(DevGeom g, const float *__restrict__ src,
                                                    float *__restrict__ dst, int ncp, int vec, int face,
                                                    int filter, int chunk) {
    const int nl = g.nlev;
    int cb, lb;
    if (!xcd_block((g.ncol + 255) / 256, (nl + chunk - 1) / chunk, cb, lb)) return;
    const int col = cb * 256 + threadIdx.x;
    if (col >= g.ncol) return;
    const int64_t nC = g.nC, nc = g.ncol;
    const int k0 = lb * chunk;
    const int k1 = min(k0 + chunk, nl);
    const int kf1 = (k1 == nl) ? nl + 1 : k1;  // the last chunk also owns face nlev
    const int m = g.gm ? g.gm[col] : 1;
    if (m <= 1) {
#pragma unroll 4
        for (int k = k0; k < kf1; ++k) {
            const int64_t q = (int64_t)k * nc + col;
            if (k < nl)
                for (int p = 0; p < ncp; ++p) dst[p * nC + q] = src[p * nC + q];
            if (face) {
                float v = src[ncp * nC + q];
                if (filter && (k == 0 || k == nl)) v = 0.0f;
                dst[ncp * nC + q] = v;
            }
        }
        return;
    }
    int mc[4];
    float mw[4], M[4][4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        mc[t] = g.gidx[t * nc + col];
        mw[t] = g.gw[t * nc + col];
#pragma unroll
        for (int e = 0; e < 4; ++e) M[t][e] = vec >= 0 ? g.gM[(4 * t + e) * nc + col] : 0.0f;
    }
    const int rank = g.grank[col];
#pragma unroll 2
    for (int k = k0; k < kf1; ++k) {
        const int64_t row = (int64_t)k * nc;
        if (k < nl) {
            for (int p = 0; p < ncp; ++p) {
                if (p == vec) {
                    float o1 = 0.0f, o2 = 0.0f;
#pragma unroll
                    for (int t = 0; t < 4; ++t) {
                        if (t >= m) break;
                        const float a = src[p * nC + row + mc[t]];
                        const float b = src[(p + 1) * nC + row + mc[t]];
                        o1 = fmaf(M[t][0], a, fmaf(M[t][1], b, o1));
                        o2 = fmaf(M[t][2], a, fmaf(M[t][3], b, o2));
                    }
                    dst[p * nC + row + col] = o1;
                    dst[(p + 1) * nC + row + col] = o2;
                    ++p;
                    continue;
                }
                float v[4];
#pragma unroll
                for (int t = 0; t < 4; ++t) v[t] = t < m ? src[p * nC + row + mc[t]] : 0.0f;
                dst[p * nC + row + col] = ranked_sum(v, mw, rank, m);
            }
        }
        if (face) {
            float s = 0.0f;
            if (!(filter && (k == 0 || k == nl))) {
                float v[4];
#pragma unroll
                for (int t = 0; t < 4; ++t) v[t] = t < m ? src[ncp * nC + row + mc[t]] : 0.0f;
                s = ranked_sum(v, mw, rank, m);
            }
            dst[ncp * nC + row + col] = s;
        }
    }
}
