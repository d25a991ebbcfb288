// ca_dss.hip -- weighted direct stiffness summation (ClimaCore
// Spaces.weighted_dss!, called at src/prognostic_equations/constrain_state.jl:59-64
// and remaining_tendency.jl:18-21).  Each shared node's value becomes the
// J-weighted average over its copies; covariant horizontal vectors are
// averaged as Cartesian vectors and projected back onto each element's basis.
//
// Gather form, out of place: thread per (column, chunk of levels).  A column
// reads every copy of its node (at most four on the cubed sphere) and writes
// only itself.  The four slots of a column's table row list the copies in
// the group's canonical order, padded with zero-weight slots, so
//   * every copy of a node sums the same values in the same order with the
//     same weights: duplicated columns receive bitwise-identical scalars;
//   * an interior node (one slot of weight 1) passes through exactly;
//   * control flow is uniform: each plane and level is one store of all
//     lanes (whole cache lines); padded slots load the column's own value
//     (always in bounds, an L1 hit) and contribute 0 * v = 0.
// The table row is read once per thread (int4 + float4 [+ 4 float4]) and
// reused over the thread's levels.
#include "ca_kernels.h"

// The 1-D grid covers (column block, level chunk) pairs, level chunk major.
// Blocks are dispatched round-robin over the 8 XCDs; the logical order is
// remapped to give every XCD a contiguous span of logical blocks: column
// blocks of one level chunk, in order, so neighbouring elements, whose
// copies a gather reads, are fetched at the same levels through that XCD's
// L2 at about the same time.  The host pads the grid to a multiple of 8
// blocks: b -> (b % 8) * q + b / 8 is then a bijection onto [0, 8 q), and
// logical blocks past the end return.
__device__ __forceinline__ bool xcd_block(int nb_col, int nb_lev, int &cb, int &lb) {
    const int nb = nb_col * nb_lev;
    const int q = (nb + 7) / 8;
    const int b = blockIdx.x;
    const int logical = (b & 7) * q + (b >> 3);
    if (logical >= nb) return false;
    lb = logical / nb_col;
    cb = logical - lb * nb_col;
    return true;
}

// The LDS-staged kernel's grid: every XCD takes a contiguous range of
// ceil(ncl / 8) element clusters, a compact patch of the sphere, for every
// level chunk, in level-chunk-major order.  The workgroups in flight on one
// XCD are then spatial neighbours at the same levels, so the element rows
// one cluster stages as halo are, mostly, rows its neighbour staged
// moments before through the same L2; only the patch's perimeter is read
// from another XCD's side.  The host launches 8 ceil(ncl / 8) nb_lev
// blocks; those mapped past the last cluster return.
__device__ __forceinline__ bool xcd_cluster_block(int ncl, int nb_lev, int &cl, int &lb) {
    const int cpx = (ncl + 7) >> 3;
    const int b = blockIdx.x;
    const int j = b >> 3;
    lb = j / cpx;
    cl = (b & 7) * cpx + (j - lb * cpx);
    return lb < nb_lev && cl < ncl;
}

// weighted sum over the four slots of one plane row, in slot order
__device__ __forceinline__ float dss_scalar_row(const float *__restrict__ row, const int mc[4], const float mw[4]) {
    float v[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) v[t] = row[mc[t]];
    float s = 0.0f;
#pragma unroll
    for (int t = 0; t < 4; ++t) s = fmaf(mw[t], v[t], s);
    return s;
}

template <int NC, int VEC, bool FACE>
__global__ void __launch_bounds__(256) k_dss_gather(DevGeom g, const float *__restrict__ src,
                                                    float *__restrict__ dst, int filter, int chunk) {
    const int nl = g.nlev;
    int cb, lb;
    if (!xcd_block((g.ncol + 255) / 256, (nl + chunk - 1) / chunk, cb, lb)) return;
    const int col = cb * 256 + threadIdx.x;
    if (col >= g.ncol) return;
    const int64_t nC = g.nC, nc = g.ncol;
    const int k0 = lb * chunk;
    const int k1 = min(k0 + chunk, nl);
    const int kf1 = (FACE && k1 == nl) ? nl + 1 : k1;  // the last chunk also owns face nlev
    const int4 ci = g.gtc[col];
    const float4 wv = g.gtw[col];
    const int mc[4] = {ci.x, ci.y, ci.z, ci.w};
    const float mw[4] = {wv.x, wv.y, wv.z, wv.w};
    float4 M[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) M[t] = VEC >= 0 ? g.gtM[4 * (int64_t)col + t] : make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll 2
    for (int k = k0; k < kf1; ++k) {
        const int64_t row = (int64_t)k * nc;
        if (k < k1) {
#pragma unroll
            for (int p = 0; p < NC; ++p) {
                if (p == VEC + 1) continue;
                const float *rp = src + p * nC + row;
                if (p == VEC) {
                    const float *rq = rp + nC;
                    float a[4], b[4];
#pragma unroll
                    for (int t = 0; t < 4; ++t) {
                        a[t] = rp[mc[t]];
                        b[t] = rq[mc[t]];
                    }
                    float o1 = 0.0f, o2 = 0.0f;
#pragma unroll
                    for (int t = 0; t < 4; ++t) {
                        o1 = fmaf(M[t].x, a[t], fmaf(M[t].y, b[t], o1));
                        o2 = fmaf(M[t].z, a[t], fmaf(M[t].w, b[t], o2));
                    }
                    dst[p * nC + row + col] = o1;
                    dst[(p + 1) * nC + row + col] = o2;
                } else {
                    dst[p * nC + row + col] = dss_scalar_row(rp, mc, mw);
                }
            }
        }
        if (FACE) {
            float s = dss_scalar_row(src + NC * nC + row, mc, mw);
            if (filter && (k == 0 || k == nl)) s = 0.0f;
            dst[NC * nC + row + col] = s;
        }
    }
}

template __global__ void k_dss_gather<4, 1, true>(DevGeom, const float *, float *, int, int);
template __global__ void k_dss_gather<4, 0, false>(DevGeom, const float *, float *, int, int);

// LDS-staged form: workgroup per (element cluster, chunk of levels).  For
// each level the workgroup stages, for every element of the cluster's list
// (own elements and the neighbours their nodes' copies live in), the 16
// node values of every plane into LDS with whole-element (64-byte) loads,
// then each thread gathers its column's copies from LDS.  A 64-lane load
// touches 4 cache lines instead of the ~30 a direct gather touches, which
// bounds the direct form by L1 tag lookups.  Two LDS buffers alternate
// between levels, so one barrier per level orders staging and reading, and
// the next level's values are loaded into registers while the current level
// is staged and used.
// Sums run over the same canonical slots with the same weights as the
// direct form: the results are bitwise identical to k_dss_gather's.
// V4: the staged element rows are loaded 4 nodes at a time (16-byte loads
// and LDS writes; src 16-byte aligned, which the host checks).
template <int NC, int VEC, bool FACE, bool V4>
__global__ void __launch_bounds__(256) k_dss_lds(DevGeom g, const float *__restrict__ src,
                                                 float *__restrict__ dst, int filter, int chunk) {
    constexpr int NP = NC + (FACE ? 1 : 0);
    constexpr int SE = CA_DSS_EMAX * 16;
    __shared__ float sh[2][NP][SE];
    const int nl = g.nlev;
    int cl, lb;
    if (!xcd_cluster_block(g.ncl, (nl + chunk - 1) / chunk, cl, lb)) return;  // uniform per workgroup
    const int tid = threadIdx.x;
    const int64_t nC = g.nC, nc = g.ncol;
    const int nS = 16 * g.dl_nE[cl];
    const int *list = g.dl_list + (int64_t)cl * CA_DSS_EMAX;
    const int64_t r = (int64_t)cl * 256 + tid;
    const int col = g.dl_col[r];
    const ushort4 sl4 = g.dl_slot[r];
    const int sl[4] = {sl4.x, sl4.y, sl4.z, sl4.w};
    const float4 wv = g.dl_w[r];
    const float mw[4] = {wv.x, wv.y, wv.z, wv.w};
    float4 M[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) M[t] = VEC >= 0 ? g.dl_M[4 * r + t] : make_float4(0.f, 0.f, 0.f, 0.f);
    const int k0 = lb * chunk;
    const int k1 = min(k0 + chunk, nl);
    const int kf1 = (FACE && k1 == nl) ? nl + 1 : k1;
    // staged values of this thread: entries tid + 256 j, j < 4, of every plane
    // (V4: the 4-node groups tid + 256 j of every plane)
    constexpr int NJ = V4 ? (CA_DSS_EMAX * 4 + 255) / 256 : (CA_DSS_EMAX * 16 + 255) / 256;
    constexpr int W = V4 ? 4 : 1;
    // byte offsets of this thread's staged entries in a level row (-1: none)
    int32_t soff[NJ];
#pragma unroll
    for (int j = 0; j < NJ; ++j) {
        const int i = tid + 256 * j;
        if (V4)
            soff[j] = i < nS / 4 ? 4 * (16 * list[i >> 2] + 4 * (i & 3)) : -1;
        else
            soff[j] = i < nS ? 4 * (16 * list[i >> 4] + (i & 15)) : -1;
    }
    float pre[NJ][NP][W];
    auto fetch = [&](int k) {
        const int64_t row = (int64_t)k * nc;
#pragma unroll
        for (int j = 0; j < NJ; ++j) {
            if (soff[j] < 0) continue;
#pragma unroll
            for (int p = 0; p < NP; ++p) {
                if (p < NC && k >= nl) {
#pragma unroll
                    for (int w = 0; w < W; ++w) pre[j][p][w] = 0.0f;
                    continue;
                }
                const char *a = reinterpret_cast<const char *>(src + p * nC + row) + (uint32_t)soff[j];
                if constexpr (V4) {
                    const float4 v = *reinterpret_cast<const float4 *>(a);
                    pre[j][p][0] = v.x;
                    pre[j][p][1] = v.y;
                    pre[j][p][2] = v.z;
                    pre[j][p][3] = v.w;
                } else {
                    pre[j][p][0] = *reinterpret_cast<const float *>(a);
                }
            }
        }
    };
    // level k + 1 is loaded into registers while level k is staged and used
    if (k0 < kf1) fetch(k0);
    int buf = 0;
    for (int k = k0; k < kf1; ++k) {
#pragma unroll
        for (int j = 0; j < NJ; ++j) {
            if (soff[j] < 0) continue;
#pragma unroll
            for (int p = 0; p < NP; ++p) {
                if constexpr (V4)
                    *reinterpret_cast<float4 *>(&sh[buf][p][4 * (tid + 256 * j)]) =
                        make_float4(pre[j][p][0], pre[j][p][1], pre[j][p][2], pre[j][p][3]);
                else
                    sh[buf][p][tid + 256 * j] = pre[j][p][0];
            }
        }
        if (k + 1 < kf1) fetch(k + 1);
        __syncthreads();
        const int64_t row = (int64_t)k * nc;
        if (col >= 0) {
            const uint32_t cb = (uint32_t)col * 4u;
            if (k < nl) {
#pragma unroll
                for (int p = 0; p < NC; ++p) {
                    if (p == VEC + 1) continue;
                    if (p == VEC) {
                        float o1 = 0.0f, o2 = 0.0f;
#pragma unroll
                        for (int t = 0; t < 4; ++t) {
                            const float a = sh[buf][p][sl[t]], b = sh[buf][p + 1][sl[t]];
                            o1 = fmaf(M[t].x, a, fmaf(M[t].y, b, o1));
                            o2 = fmaf(M[t].z, a, fmaf(M[t].w, b, o2));
                        }
                        ca_sto(dst + p * nC + row, cb, o1);
                        ca_sto(dst + (p + 1) * nC + row, cb, o2);
                    } else {
                        float s = 0.0f;
#pragma unroll
                        for (int t = 0; t < 4; ++t) s = fmaf(mw[t], sh[buf][p][sl[t]], s);
                        ca_sto(dst + p * nC + row, cb, s);
                    }
                }
            }
            if (FACE) {
                float s = 0.0f;
#pragma unroll
                for (int t = 0; t < 4; ++t) s = fmaf(mw[t], sh[buf][NC][sl[t]], s);
                if (filter && (k == 0 || k == nl)) s = 0.0f;
                ca_sto(dst + NC * nC + row, cb, s);
            }
        }
        buf ^= 1;
    }
}

template __global__ void k_dss_lds<4, 1, true, false>(DevGeom, const float *, float *, int, int);
template __global__ void k_dss_lds<4, 0, false, false>(DevGeom, const float *, float *, int, int);
template __global__ void k_dss_lds<4, 1, true, true>(DevGeom, const float *, float *, int, int);
template __global__ void k_dss_lds<4, 0, false, true>(DevGeom, const float *, float *, int, int);
