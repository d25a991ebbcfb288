// ca_newton_pcr.hip -- the fused implicit ARS343 stage with one wavefront per
// column and one lane per level (nlev + 1 <= 64 faces).
//
// Same algorithm as k_newton_stage (ca_column.hip): stage-state assembly,
// impenetrability filter, one Newton iteration with the manual Jacobian of
// manual_sparse_jacobian.jl:713-870, Schur complement on u_3
// (manual_sparse_jacobian.jl:579-584), post-Newton upwind correction
// (implicit_tendency.jl:322-339), and F_I = (U - U~) / dtg.  The tridiagonal
// Schur system is solved by parallel cyclic reduction across the 64 lanes.
//
// A workgroup owns a tile of TC adjacent columns: the tile is assembled from
// the sources with coalesced row loads into LDS, each wavefront then solves
// its columns lane-per-level from LDS, and the results leave through LDS
// with coalesced row stores.
#include "ca_kernels.h"

namespace pcr {

template <class T>
__device__ __forceinline__ T up(T v, int d) { return __shfl_up(v, d, 64); }
template <class T>
__device__ __forceinline__ T dn(T v, int d) { return __shfl_down(v, d, 64); }

// Shifts by one lane across the whole wavefront as DPP wave shifts
// (wave_shr:1, lane i <- lane i - 1; wave_shl:1, lane i <- lane i + 1) in the
// operand of a VALU move instead of an LDS-crossbar permute.  The edge lane
// has no source and keeps its own value (old = v, bound_ctrl off), as
// __shfl_up / __shfl_down leave an out-of-range lane.  All lanes are active
// wherever these are called (uniform control flow).
__device__ __forceinline__ int up1i(int v) { return __builtin_amdgcn_update_dpp(v, v, 0x138, 0xF, 0xF, false); }
__device__ __forceinline__ int dn1i(int v) { return __builtin_amdgcn_update_dpp(v, v, 0x130, 0xF, 0xF, false); }
__device__ __forceinline__ float up1(float v) { return __int_as_float(up1i(__float_as_int(v))); }
__device__ __forceinline__ float dn1(float v) { return __int_as_float(dn1i(__float_as_int(v))); }
__device__ __forceinline__ double up1(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = up1i((int)(b & 0xffffffffll)), hi = up1i((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ double dn1(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = dn1i((int)(b & 0xffffffffll)), hi = dn1i((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// lane 0's value in every lane (v_readlane_b32 into a scalar register)
__device__ __forceinline__ float lane0(float v) { return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 0)); }

// Both neighbours at distance 32 from one v_permlane32_swap of v with itself:
// the swap exchanges lanes 32-63 of its first operand with lanes 0-31 of its
// second, so the first result holds lane i - 32 in lanes 32-63 (own value
// below) and the second lane i + 32 in lanes 0-31 (own value above): exactly
// __shfl_up(v, 32) and __shfl_down(v, 32).
__device__ __forceinline__ void swap32(double v, double &u, double &d) {
    const long long b = __double_as_longlong(v);
    const unsigned lo = (unsigned)(b & 0xffffffffll), hi = (unsigned)(b >> 32);
    const auto l = __builtin_amdgcn_permlane32_swap(lo, lo, false, false);
    const auto h = __builtin_amdgcn_permlane32_swap(hi, hi, false, false);
    u = __longlong_as_double(((long long)h[0] << 32) | l[0]);
    d = __longlong_as_double(((long long)h[1] << 32) | l[1]);
}

__device__ __forceinline__ void swap32(float v, float &u, float &d) {
    const unsigned b = __float_as_uint(v);
    const auto r = __builtin_amdgcn_permlane32_swap(b, b, false, false);
    u = __uint_as_float(r[0]);
    d = __uint_as_float(r[1]);
}

// The rows k - s (u) and k + s (d) of a parallel-cyclic-reduction step, s a
// power of two below 64 that is a constant once the step loop is unrolled.
// s = 1: DPP wave shifts; s = 2: two of them; s = 32: one lane swap per
// dword; otherwise LDS-crossbar permutes.  Out-of-range lanes receive some
// lane's finite value (with s = 2 lane 1 receives lane 0's, not its own):
// PCR multiplies every such value by an exactly zero coupling.
template <class T>
__device__ __forceinline__ void pcr_nbrs(T v, int s, T &u, T &d) {
    if (s == 1) {
        u = up1(v);
        d = dn1(v);
    } else if (s == 2) {
        u = up1(up1(v));
        d = dn1(dn1(v));
    } else if (s == 32) {
        swap32(v, u, d);
    } else {
        u = up(v, s);
        d = dn(v, s);
    }
}

}  // namespace pcr

// NF / NI: compile-time numbers of full-state and implicit-tendency sources
// (the three ARS343 implicit stages have (2, 0), (3, 1), (4, 2)); -1 reads
// the counts from src.
template <int NF, int NI>
__device__ __forceinline__ void assemble(const StageSrc &src, int64_t nC, int64_t q, bool okc, bool cen,
                                         bool face, float out[5]) {
    const int nf = NF >= 0 ? NF : src.nfull;
    const int ni = NI >= 0 ? NI : src.nfi;
#pragma unroll
    for (int comp = 0; comp < 4; ++comp) {
        float s = 0.0f;
        if (okc && cen) {
#pragma unroll
            for (int t = 0; t < (NF >= 0 ? NF : CA_MAX_SRC); ++t)
                if (t < nf) s = fmaf(src.cf[t], src.full[t][comp * nC + q], s);
            if (comp == 0 || comp == 3) {
                const int ci = comp == 0 ? 0 : 1;
#pragma unroll
                for (int t = 0; t < (NI >= 0 ? NI : CA_MAX_SRC); ++t)
                    if (t < ni) s = fmaf(src.ci[t], src.fi[t][ci * nC + q], s);
            }
        }
        out[comp] = s;
    }
    float s = 0.0f;
    if (okc && face) {
#pragma unroll
        for (int t = 0; t < (NF >= 0 ? NF : CA_MAX_SRC); ++t)
            if (t < nf) s = fmaf(src.cf[t], src.full[t][4 * nC + q], s);
#pragma unroll
        for (int t = 0; t < (NI >= 0 ? NI : CA_MAX_SRC); ++t)
            if (t < ni) s = fmaf(src.ci[t], src.fi[t][2 * nC + q], s);
    }
    out[4] = s;
}

// As assemble, for the 4 adjacent columns q .. q + 3 with 16-byte loads
// (q, nC and the source pointers 16-byte aligned); the same fmaf order per
// column, so the sums are bitwise those of assemble.
__device__ __forceinline__ float4 ca_ld4(const float *p) { return *reinterpret_cast<const float4 *>(p); }
__device__ __forceinline__ void ca_fma4(float a, const float4 &x, float4 &s) {
    s.x = fmaf(a, x.x, s.x);
    s.y = fmaf(a, x.y, s.y);
    s.z = fmaf(a, x.z, s.z);
    s.w = fmaf(a, x.w, s.w);
}
template <int NF, int NI>
__device__ __forceinline__ void assemble4(const StageSrc &src, int64_t nC, int64_t q, bool okc, bool cen,
                                          bool face, float4 out[5]) {
    const int nf = NF >= 0 ? NF : src.nfull;
    const int ni = NI >= 0 ? NI : src.nfi;
#pragma unroll
    for (int comp = 0; comp < 4; ++comp) {
        float4 s = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (okc && cen) {
#pragma unroll
            for (int t = 0; t < (NF >= 0 ? NF : CA_MAX_SRC); ++t)
                if (t < nf) ca_fma4(src.cf[t], ca_ld4(src.full[t] + comp * nC + q), s);
            if (comp == 0 || comp == 3) {
                const int ci = comp == 0 ? 0 : 1;
#pragma unroll
                for (int t = 0; t < (NI >= 0 ? NI : CA_MAX_SRC); ++t)
                    if (t < ni) ca_fma4(src.ci[t], ca_ld4(src.fi[t] + ci * nC + q), s);
            }
        }
        out[comp] = s;
    }
    float4 s = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (okc && face) {
#pragma unroll
        for (int t = 0; t < (NF >= 0 ? NF : CA_MAX_SRC); ++t)
            if (t < nf) ca_fma4(src.cf[t], ca_ld4(src.full[t] + 4 * nC + q), s);
#pragma unroll
        for (int t = 0; t < (NI >= 0 ? NI : CA_MAX_SRC); ++t)
            if (t < ni) ca_fma4(src.ci[t], ca_ld4(src.fi[t] + 2 * nC + q), s);
    }
    out[4] = s;
}

// V4: 16-byte row accesses in phases 1 and 3 (ncol a multiple of 4 and every
// source and destination 16-byte aligned; the host checks).
template <int TC, int NF, int NI, bool V4>
__global__ void __launch_bounds__(256) k_newton_pcr(DevGeom g, DevNum n, StageSrc src, float dtg, float *U,
                                                    float *FI, float *Ut) {
    constexpr int R = TC + 1;  // padded LDS row (conflict-free column reads)
    __shared__ float tU[5][64][R];
    __shared__ float tF[3][64][R];
    const int nl = g.nlev;
    const int64_t nC = g.nC, nc = g.ncol;
    const int col0 = ca_xcd_tile(blockIdx.x, gridDim.x) * TC;
    const int tid = threadIdx.x;

    // ---- phase 1: assemble U~ (coalesced rows of TC columns)
    if constexpr (V4) {
        for (int idx = tid; idx < 16 * TC; idx += 256) {
            const int r = idx / (TC / 4), c4 = 4 * (idx - r * (TC / 4));
            const int col = col0 + c4;
            const int64_t q = (int64_t)r * nc + col;
            float4 v[5];
            assemble4<NF, NI>(src, nC, q, col < nc, r < nl, r >= 1 && r <= nl - 1, v);
#pragma unroll
            for (int comp = 0; comp < 5; ++comp) {
                tU[comp][r][c4] = v[comp].x;
                tU[comp][r][c4 + 1] = v[comp].y;
                tU[comp][r][c4 + 2] = v[comp].z;
                tU[comp][r][c4 + 3] = v[comp].w;
            }
            if (Ut && col < nc) {
                if (r < nl)
#pragma unroll
                    for (int comp = 0; comp < 4; ++comp) *reinterpret_cast<float4 *>(Ut + comp * nC + q) = v[comp];
                if (r <= nl) *reinterpret_cast<float4 *>(Ut + 4 * nC + q) = v[4];
            }
        }
    } else {
#pragma unroll 2
    for (int idx = tid; idx < 64 * TC; idx += 256) {
        const int r = idx / TC, cc = idx - r * TC;
        const int col = col0 + cc;
        float v[5];
        assemble<NF, NI>(src, nC, (int64_t)r * nc + col, col < nc, r < nl, r >= 1 && r <= nl - 1, v);
#pragma unroll
        for (int comp = 0; comp < 5; ++comp) tU[comp][r][cc] = v[comp];
        // U~ for the reverse pass (when recording), in the state layout
        if (Ut && col < nc) {
            const int64_t q = (int64_t)r * nc + col;
            if (r < nl)
#pragma unroll
                for (int comp = 0; comp < 4; ++comp) Ut[comp * nC + q] = v[comp];
            if (r <= nl) Ut[4 * nC + q] = v[4];
        }
    }
    }
    __syncthreads();

    // ---- phase 2: one column per wavefront, lane k = centre k and face k
    const int k = tid & 63;
    const int wid = tid >> 6;
    const bool hasC = k < nl;
    const bool intF = k >= 1 && k <= nl - 1;
    const bool isF = k <= nl;
    const int kc = hasC ? k : nl - 1;   // clamped centre level for padding lanes
    const int kf = isF ? k : nl;        // clamped face level
    const float g33 = g.g33f[kf];
    const float g33p = g.g33f[kc + 1];
    const float idtg = 1.0f / dtg;
    for (int cc = wid; cc < TC; cc += 4) {
        const int col = col0 + cc;
        if (col >= nc) break;  // uniform across the wavefront
        const float rho = tU[0][kc][cc], u1 = tU[1][kc][cc], u2 = tU[2][kc][cc], rhoe = tU[3][kc][cc];
        const float u3 = intF ? tU[4][k][cc] : 0.0f;
        const float u3p = pcr::dn1(u3);  // face k + 1
        const float u3m = pcr::up1(u3);  // face k - 1
        const CV c = cv_make(g, n, kc, col, rho, u1, u2, rhoe, u3, hasC ? u3p : 0.0f);
        const float rJ = c.rho * c.J;
        // centre row k: A(rho,u3), A(rhoe,u3) on faces k, k+1 and R = dtg T_imp
        const float rJm = pcr::up1(rJ), hm = pcr::up1(c.h);
        const float rJp = pcr::dn1(rJ), hp = pcr::dn1(c.h);
        float a0 = 0.0f, a1 = 0.0f, e0 = 0.0f, e1 = 0.0f, Rr = 0.0f, Re = 0.0f;
        if (hasC) {
            float Fm = 0.0f, Fp = 0.0f, Hm = 0.0f, Hp = 0.0f;
            if (k >= 1) {
                const float rJf = 0.5f * (rJm + rJ);
                a0 = dtg * (rJf * c.iJ) * g33;
                e0 = a0 * (0.5f * (hm + c.h));
                Fm = rJf * (u3 * g33);
                Hm = Fm * (0.5f * (hm + c.h));
            }
            if (k + 1 <= nl - 1) {
                const float rJf = 0.5f * (rJ + rJp);
                a1 = dtg * (-rJf * c.iJ) * g33p;
                e1 = a1 * (0.5f * (c.h + hp));
                Fp = rJf * (u3p * g33p);
                Hp = Fp * (0.5f * (c.h + hp));
            }
            Rr = dtg * (-(Fp - Fm) * c.iJ);
            Re = dtg * (-(Hp - Hm) * c.iJ);
        }
        // face row k: needs centre k - 1 (lane k - 1) and centre k (own lane)
        CV a;
        a.rho = pcr::up1(c.rho);
        a.K = pcr::up1(c.K);
        a.Td = pcr::up1(c.Td);
        a.pd = pcr::up1(c.pd);
        a.Pid = pcr::up1(c.Pid);
        a.T = (float)a.Td;
        a.p = (float)a.pd;
        a.Pi = (float)a.Pid;
        a.ct1 = pcr::up1(c.ct1);
        a.ct2 = pcr::up1(c.ct2);
        const float pa0 = pcr::up1(a0), pa1 = pcr::up1(a1), pe0 = pcr::up1(e0), pe1 = pcr::up1(e1);
        const float pRr = pcr::up1(Rr), pRe = pcr::up1(Re);
        float L = 0.0f, D = 1.0f, Up = 0.0f, rhs = 0.0f;
        if (intF) {
            const FaceRow fr = jac_face_row(g, n, k, col, a, c, u3m, u3, u3p, dtg);
            const float Ru3 = dtg * (pgf_u3(g, n, k, a, c) - beta_w(g, n, k, col) * u3);
            L = fr.lo + fr.r0 * pa0 + fr.e0 * pe0;
            D = fr.d + fr.r0 * pa1 + fr.e0 * pe1 + fr.r1 * a0 + fr.e1 * e0;
            Up = fr.up + fr.r1 * a1 + fr.e1 * e1;
            rhs = Ru3 + fr.r0 * pRr + fr.r1 * Rr + fr.e0 * pRe + fr.e1 * Re;
        } else if (isF) {
            D = -dtg * beta_w(g, n, k, col) - 1.0f;
        }
        // parallel cyclic reduction on the 64 face equations, in float64: the
        // acoustic Schur system is stiff (dtg c / dz >> 1) and PCR's
        // elimination order loses more in float32 than a Thomas sweep.
        // Step s eliminates the couplings to rows k -+ s with the neighbour
        // rows divided by their own diagonal: each lane shares
        // (L, U, r) / D and forms
        //   L' = -L La~,  U' = -U Ub~,  D' = D - L Ua~ - U Lb~,  r' = r - L ra~ - U rb~.
        // Row k couples only to rows k -+ s after step s / 2, so for k < s
        // (k + s > 63) its L (U) is exactly 0, and the out-of-range
        // exchanges, which return some lane's finite values, are multiplied
        // by 0 (pcr::pcr_nbrs).
        double Ld = L, Dd = D, Ud = Up, rd = rhs;
#pragma unroll
        for (int s = 1; s < 64; s <<= 1) {
            const double iD = ca_drcp(Dd);
            const double Ln = Ld * iD, Un = Ud * iD, rn = rd * iD;
            double La, Ua, ra, Lb, Ub, rb;
            pcr::pcr_nbrs(Ln, s, La, Lb);
            pcr::pcr_nbrs(Un, s, Ua, Ub);
            pcr::pcr_nbrs(rn, s, ra, rb);
            const double nD = Dd - Ld * Ua - Ud * Lb;
            const double nr = rd - Ld * ra - Ud * rb;
            Ld = -Ld * La;
            Ud = -Ud * Ub;
            Dd = nD;
            rd = nr;
        }
        const float x = (float)(rd * ca_drcp(Dd));  // du_3 at face k
        const float x1 = pcr::dn1(x);   // du_3 at face k + 1
        // Newton update
        const float dr = a0 * x + a1 * x1 - Rr;
        const float de = e0 * x + e1 * x1 - Re;
        const float rho_n = rho - dr;
        float rhoe_n = rhoe - de;
        const float u3_n = intF ? u3 - x : 0.0f;
        float fi_e = -de * idtg;
        // post-Newton upwind correction at the new state
        if (n.upw != 0) {
            const float u3_np = pcr::dn1(u3_n);
            const float hn = cv_enthalpy(g, n, kc, col, rho_n, u1, u2, rhoe_n, u3_n, hasC ? u3_np : 0.0f);
            const float href = pcr::lane0(hn);
            const float hr = hn - href;
            const float sk = g.sc[kc];
            const float rJn = rho_n * (g.Jh[col] * sk * sk * g.Jvc[kc]);  // J: geometry only, reloaded
            const float h2 = pcr::up(hr, 2), h1 = pcr::up1(hr), h4 = pcr::dn1(hr);
            const float rJn1 = pcr::up1(rJn);
            float dF = 0.0f;
            if (intF) {
                const float v = u3_n * g33;
                const float rJf = 0.5f * (rJn1 + rJn);
                const float upw = upwind_product(n.upw, k, nl, v, n.dt, h2, h1, hr, h4);
                const float cen = upwind_product(0, k, nl, v, n.dt, 0.0f, h1, hr, 0.0f);
                dF = rJf * upw - rJf * cen;
            }
            const float dF1 = pcr::dn1(dF);
            const float corr = hasC ? -(dF1 - dF) * (g.iJh[col] * g.isJc[kc]) : 0.0f;
            rhoe_n += dtg * corr;
            fi_e += corr;
        }
        if (hasC) {
            tU[0][k][cc] = rho_n;
            tU[3][k][cc] = rhoe_n;
            tF[0][k][cc] = -dr * idtg;
            tF[1][k][cc] = fi_e;
        }
        if (isF) {
            tU[4][k][cc] = u3_n;
            tF[2][k][cc] = intF ? -x * idtg : 0.0f;
        }
    }
    __syncthreads();

    // ---- phase 3: coalesced stores of U and F_I
    if constexpr (V4) {
        for (int idx = tid; idx < 16 * TC; idx += 256) {
            const int r = idx / (TC / 4), c4 = 4 * (idx - r * (TC / 4));
            const int col = col0 + c4;
            if (col >= nc) continue;
            const int64_t q = (int64_t)r * nc + col;
            auto row4 = [&](const float (*t)[R]) {
                return make_float4(t[r][c4], t[r][c4 + 1], t[r][c4 + 2], t[r][c4 + 3]);
            };
            if (r < nl) {
#pragma unroll
                for (int comp = 0; comp < 4; ++comp) *reinterpret_cast<float4 *>(U + comp * nC + q) = row4(tU[comp]);
                if (FI) {
                    *reinterpret_cast<float4 *>(FI + q) = row4(tF[0]);
                    *reinterpret_cast<float4 *>(FI + nC + q) = row4(tF[1]);
                }
            }
            if (r <= nl) {
                *reinterpret_cast<float4 *>(U + 4 * nC + q) = row4(tU[4]);
                if (FI) *reinterpret_cast<float4 *>(FI + 2 * nC + q) = row4(tF[2]);
            }
        }
        return;
    }
    for (int idx = tid; idx < 64 * TC; idx += 256) {
        const int r = idx / TC, cc = idx - r * TC;
        const int col = col0 + cc;
        if (col >= nc) continue;
        const int64_t q = (int64_t)r * nc + col;
        if (r < nl) {
#pragma unroll
            for (int comp = 0; comp < 4; ++comp) U[comp * nC + q] = tU[comp][r][cc];
            if (FI) {
                FI[q] = tF[0][r][cc];
                FI[nC + q] = tF[1][r][cc];
            }
        }
        if (r <= nl) {
            U[4 * nC + q] = tU[4][r][cc];
            if (FI) FI[2 * nC + q] = tF[2][r][cc];
        }
    }
}

#define CA_INST_PCR(TC, NF, NI, V4)                                                                   \
    template __global__ void k_newton_pcr<TC, NF, NI, V4>(DevGeom, DevNum, StageSrc, float, float *, float *, \
                                                          float *);
CA_INST_PCR(16, 2, 0, false)
CA_INST_PCR(16, 3, 1, false)
CA_INST_PCR(16, 4, 2, false)
CA_INST_PCR(16, -1, -1, false)
CA_INST_PCR(16, 2, 0, true)
CA_INST_PCR(16, 3, 1, true)
CA_INST_PCR(16, 4, 2, true)
CA_INST_PCR(16, -1, -1, true)
CA_INST_PCR(32, 2, 0, false)
CA_INST_PCR(32, 3, 1, false)
CA_INST_PCR(32, 4, 2, false)
CA_INST_PCR(32, -1, -1, false)
