// Hexahedral operators whose modes and points differ per direction (FP64, runtime sizes).
//
// The reference's MatrixFree kernels are instantiated per (modes, points) with the same
// values in all three directions and silently use direction 0's tables
// (MatrixFreeOps/Helmholtz.h:669-760, :844-856); expansions with NUMMODES="4,6,8" run
// through its StdRegions/LocalRegions matrix path instead (e.g. the session of
// Demos/MultiRegions/Tests/Helmholtz3D_CG_Hex_Heterogeneous.tst).  These kernels apply the
// same operators for any (nm0, nm1, nm2, nq0, nq1, nq2): one workgroup per element, every
// 1D contraction a block-cooperative pass over a 3D array in LDS with the direction's table
// (staged in LDS), so one compiled kernel covers every size combination.  Geometry is read
// in the reference's own GetJac / GetDerivFactors layouts (no packing): the Helmholtz
// metric G^{ab} = sum_c df[c*3+a] df[c*3+b] (Helmholtz.h:858-966) is formed per point.
#include <hip/hip_runtime.h>

#include "imf_common.hpp"
#include "mf_aniso.hpp"

namespace imf {

namespace {

// out = M applied along `axis` of the (n0, n1, n2) array `in` (axis 0 fastest): the axis
// length becomes nout and out[.., x, ..] = sum_a in[.., a, ..] * Mt[a*sa + x*sx].
// With ACC the result is added to out.
template <bool ACC>
__device__ void apply_axis(const double *__restrict__ in, double *__restrict__ out, int n0,
                           int n1, int n2, int axis, int nout, const double *__restrict__ Mt,
                           int sa, int sx)
{
    const int o0 = axis == 0 ? nout : n0, o1 = axis == 1 ? nout : n1,
              o2 = axis == 2 ? nout : n2;
    const int nin = axis == 0 ? n0 : axis == 1 ? n1 : n2;
    const int stride = axis == 0 ? 1 : axis == 1 ? n0 : n0 * n1;
    const int tot = o0 * o1 * o2;
    for (int idx = threadIdx.x; idx < tot; idx += blockDim.x) {
        const int x0 = idx % o0, x1 = (idx / o0) % o1, x2 = idx / (o0 * o1);
        const int x = axis == 0 ? x0 : axis == 1 ? x1 : x2;
        const int base = (axis == 0 ? 0 : x0) + (axis == 1 ? 0 : x1) * n0 +
                         (axis == 2 ? 0 : x2) * n0 * n1;
        double s = 0.0;
        for (int a = 0; a < nin; ++a) s = fma(in[base + a * stride], Mt[a * sa + x * sx], s);
        if (ACC) out[idx] += s;
        else out[idx] = s;
    }
}

struct Sizes {
    int nm[3], nq[3];
    const double *B[3], *D[3], *w[3];
};

// (B2 x B1 x B0) c: modes (nm0, nm1, nm2) in a -> points (nq0, nq1, nq2) in dst (uses b).
__device__ void bwd(const Sizes &z, double *a, double *b, double *dst)
{
    apply_axis<false>(a, b, z.nm[0], z.nm[1], z.nm[2], 0, z.nq[0], z.B[0], z.nq[0], 1);
    __syncthreads();
    apply_axis<false>(b, a, z.nq[0], z.nm[1], z.nm[2], 1, z.nq[1], z.B[1], z.nq[1], 1);
    __syncthreads();
    apply_axis<false>(a, dst, z.nq[0], z.nq[1], z.nm[2], 2, z.nq[2], z.B[2], z.nq[2], 1);
    __syncthreads();
}

// (B2 x B1 x B0)^T f: points in f -> modes, written by `store(index, value)` (uses a, b).
template <typename StoreFn>
__device__ void iprod(const Sizes &z, const double *f, double *a, double *b, StoreFn store)
{
    apply_axis<false>(f, a, z.nq[0], z.nq[1], z.nq[2], 2, z.nm[2], z.B[2], 1, z.nq[2]);
    __syncthreads();
    apply_axis<false>(a, b, z.nq[0], z.nq[1], z.nm[2], 1, z.nm[1], z.B[1], 1, z.nq[1]);
    __syncthreads();
    // last contraction straight to the output (one value per thread, coalesced)
    const int tot = z.nm[0] * z.nm[1] * z.nm[2], n0 = z.nq[0];
    for (int idx = threadIdx.x; idx < tot; idx += blockDim.x) {
        const int p = idx % z.nm[0], rest = idx / z.nm[0];
        double s = 0.0;
        for (int i = 0; i < n0; ++i) s = fma(b[rest * n0 + i], z.B[0][p * n0 + i], s);
        store(idx, s);
    }
    __syncthreads();
}

}  // namespace

__global__ __launch_bounds__(256) void hex_aniso_kernel(AnisoArgs g)
{
    extern __shared__ double sm[];
    Sizes z;
    for (int d = 0; d < 3; ++d) {
        z.nm[d] = g.nm[d];
        z.nq[d] = g.nq[d];
    }
    // table image: B0 B1 B2 D0 D1 D2 w0 w1 w2 (aniso_table_len doubles), then 4 buffers
    const int tl = aniso_table_len(g.nm, g.nq);
    for (int i = threadIdx.x; i < tl; i += blockDim.x) sm[i] = g.tab[i];
    {
        int o = 0;
        for (int d = 0; d < 3; ++d) { z.B[d] = sm + o; o += z.nm[d] * z.nq[d]; }
        for (int d = 0; d < 3; ++d) { z.D[d] = sm + o; o += z.nq[d] * z.nq[d]; }
        for (int d = 0; d < 3; ++d) { z.w[d] = sm + o; o += z.nq[d]; }
    }
    double *b0 = sm + tl, *b1 = b0 + g.bufLen, *b2 = b1 + g.bufLen, *b3 = b2 + g.bufLen;
    __syncthreads();
    const long e = blockIdx.x;
    const int M = z.nm[0] * z.nm[1] * z.nm[2], Q = z.nq[0] * z.nq[1] * z.nq[2];
    const long nQ = (long)g.nElmt * Q;
    // w J at point q of element e
    auto wj = [&](int q) {
        const int i = q % z.nq[0], j = (q / z.nq[0]) % z.nq[1], k = q / (z.nq[0] * z.nq[1]);
        const double w = z.w[0][i] * z.w[1][j] * z.w[2][k];
        return w * (g.geo == kDeformed ? g.jac[e * Q + q] : g.jac[e]);
    };
    for (int v = 0; v < g.nvec; ++v) {
        const double *in = g.in + v * g.ldin;
        double *out = g.out + v * g.ldout;
        if (g.kind == kBwdTrans) {
            for (int i = threadIdx.x; i < M; i += blockDim.x) b0[i] = in[e * M + i];
            __syncthreads();
            bwd(z, b0, b1, b2);
            for (int i = threadIdx.x; i < Q; i += blockDim.x) out[e * Q + i] = b2[i];
        } else if (g.kind == kIProduct) {
            for (int q = threadIdx.x; q < Q; q += blockDim.x) b2[q] = in[e * Q + q] * wj(q);
            __syncthreads();
            iprod(z, b2, b0, b1, [&](int i, double s) { out[e * M + i] = s; });
        } else if (g.kind == kMass) {
            for (int i = threadIdx.x; i < M; i += blockDim.x) b0[i] = in[e * M + i];
            __syncthreads();
            bwd(z, b0, b1, b2);
            for (int q = threadIdx.x; q < Q; q += blockDim.x) b2[q] *= wj(q);
            __syncthreads();
            iprod(z, b2, b0, b1, [&](int i, double s) { out[e * M + i] = s; });
        } else {  // Helmholtz
            for (int i = threadIdx.x; i < M; i += blockDim.x) b0[i] = in[e * M + i];
            __syncthreads();
            bwd(z, b0, b1, b3);  // u in b3
            // reference-space derivatives d_a (PhysDerivTensor3DKernel) into b0, b1, b2
            apply_axis<false>(b3, b0, z.nq[0], z.nq[1], z.nq[2], 0, z.nq[0], z.D[0], z.nq[0], 1);
            apply_axis<false>(b3, b1, z.nq[0], z.nq[1], z.nq[2], 1, z.nq[1], z.D[1], z.nq[1], 1);
            apply_axis<false>(b3, b2, z.nq[0], z.nq[1], z.nq[2], 2, z.nq[2], z.D[2], z.nq[2], 1);
            __syncthreads();
            // g_a = w J sum_b G^{ab} d_b in place of d_a, lambda w J u in place of u
            for (int q = threadIdx.x; q < Q; q += blockDim.x) {
                double df[9];
                for (int d = 0; d < 9; ++d)
                    df[d] = g.geo == kDeformed ? g.df[d * nQ + e * Q + q] : g.df[d * g.nElmt + e];
                double G[3][3];
                for (int a = 0; a < 3; ++a)
                    for (int b = 0; b < 3; ++b)
                        G[a][b] = df[a] * df[b] + df[3 + a] * df[3 + b] + df[6 + a] * df[6 + b];
                const double jw = wj(q), d0 = b0[q], d1 = b1[q], d2 = b2[q];
                b0[q] = jw * (G[0][0] * d0 + G[0][1] * d1 + G[0][2] * d2);
                b1[q] = jw * (G[1][0] * d0 + G[1][1] * d1 + G[1][2] * d2);
                b2[q] = jw * (G[2][0] * d0 + G[2][1] * d1 + G[2][2] * d2);
                b3[q] *= g.lambda * jw;
            }
            __syncthreads();
            // h = lambda w J u + D0^T g0 + D1^T g1 + D2^T g2 (collocation transposes)
            apply_axis<true>(b0, b3, z.nq[0], z.nq[1], z.nq[2], 0, z.nq[0], z.D[0], 1, z.nq[0]);
            __syncthreads();
            apply_axis<true>(b1, b3, z.nq[0], z.nq[1], z.nq[2], 1, z.nq[1], z.D[1], 1, z.nq[1]);
            __syncthreads();
            apply_axis<true>(b2, b3, z.nq[0], z.nq[1], z.nq[2], 2, z.nq[2], z.D[2], 1, z.nq[2]);
            __syncthreads();
            iprod(z, b3, b0, b1, [&](int i, double s) { out[e * M + i] = s; });
        }
        __syncthreads();
    }
}

// Local diagonal of the anisotropic Helmholtz operator (the assembled solver's Jacobi
// preconditioner): one thread per (element, mode).  dB_d = D_d B_d per direction, taken
// from the table (B, D, w) directly.
__global__ void helm_diag_hex_aniso(AnisoArgs g, double *__restrict__ diag)
{
    const int nm0 = g.nm[0], nm1 = g.nm[1], nm2 = g.nm[2];
    const int nq0 = g.nq[0], nq1 = g.nq[1], nq2 = g.nq[2];
    const long M = (long)nm0 * nm1 * nm2, Q = (long)nq0 * nq1 * nq2;
    const long gi = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (gi >= (long)g.nElmt * M) return;
    const long e = gi / M;
    const int md = (int)(gi - e * M), p = md % nm0, q = (md / nm0) % nm1, r = md / (nm0 * nm1);
    const double *B0 = g.tab, *B1 = B0 + nm0 * nq0, *B2 = B1 + nm1 * nq1;
    const double *D0 = B2 + nm2 * nq2, *D1 = D0 + nq0 * nq0, *D2 = D1 + nq1 * nq1;
    const double *w0 = D2 + nq2 * nq2, *w1 = w0 + nq0, *w2 = w1 + nq1;
    auto dB = [](const double *B, const double *D, int nq, int m, int x) {
        double s = 0.0;
        for (int l = 0; l < nq; ++l) s += D[l * nq + x] * B[m * nq + l];
        return s;
    };
    const long nQ = (long)g.nElmt * Q;
    double s = 0.0;
    for (int k = 0; k < nq2; ++k)
        for (int j = 0; j < nq1; ++j)
            for (int i = 0; i < nq0; ++i) {
                const long pt = ((long)k * nq1 + j) * nq0 + i;
                const double bp = B0[p * nq0 + i], bq = B1[q * nq1 + j], br = B2[r * nq2 + k];
                const double phi = bp * bq * br;
                const double dd[3] = {dB(B0, D0, nq0, p, i) * bq * br, bp * dB(B1, D1, nq1, q, j) * br,
                                      bp * bq * dB(B2, D2, nq2, r, k)};
                double df[9];
                for (int d = 0; d < 9; ++d)
                    df[d] = g.geo == kDeformed ? g.df[d * nQ + e * Q + pt] : g.df[d * g.nElmt + e];
                const double jw = w0[i] * w1[j] * w2[k] * (g.geo == kDeformed ? g.jac[e * Q + pt] : g.jac[e]);
                double lap = 0.0;
                for (int a = 0; a < 3; ++a)
                    for (int b = 0; b < 3; ++b)
                        lap += (df[a] * df[b] + df[3 + a] * df[3 + b] + df[6 + a] * df[6 + b]) * dd[a] * dd[b];
                s += jw * (g.lambda * phi * phi + lap);
            }
    diag[gi] = s;
}

hipError_t launch_aniso(const AnisoArgs &a, hipStream_t s)
{
    const size_t lds = sizeof(double) * (aniso_table_len(a.nm, a.nq) + 4 * (size_t)a.bufLen);
    hipError_t e = hipFuncSetAttribute((const void *)&hex_aniso_kernel,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    hex_aniso_kernel<<<a.nElmt, 256, lds, s>>>(a);
    return hipGetLastError();
}

hipError_t launch_aniso_diag(const AnisoArgs &a, double *diag, hipStream_t s)
{
    const long n = (long)a.nElmt * a.nm[0] * a.nm[1] * a.nm[2];
    helm_diag_hex_aniso<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(a, diag);
    return hipGetLastError();
}

}  // namespace imf
