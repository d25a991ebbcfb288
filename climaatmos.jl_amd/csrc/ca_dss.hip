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
