// Quadrilateral sum-factorised operators for gfx950.
//
// One element is owned by NQ threads, thread i holding the j-line of points (i, 0..NQ-1)
// in registers; contractions along j run in registers, contractions along i run on
// i-lines re-oriented through LDS.  Elements are packed 4*floor(64/NQ) to a 256-thread
// block.  Semantics follow (paths relative to the reference root):
//   BwdTransQuadKernel   BwdTransKernels.hpp:35-76
//   IProductQuadKernel   IProductKernels.hpp:76-133   (stores out[q*nm + p])
//   PhysDerivTensor2D    PhysDerivKernels.hpp:39-91
//   HelmholtzQuadImpl    Helmholtz.h:138-276
#pragma once

#include "imf_common.hpp"
#include "mf_hex_kernels.hpp"  // ld_stream / st_stream

namespace imf {

// u_hat (global) -> u (registers: thread i holds u[j] at point (i, j)).  Uses buf0.
template <int NM, int NQ, typename T>
__device__ __forceinline__ void quad_bwd_to_lines(const Tab<NM, NQ, T> &tab,
                                                  const T *__restrict__ src, bool valid,
                                                  bool active, int t, T *buf0, T (&u)[NQ])
{
    if (active && t < NM) {
        T uh[NM];
        if (valid) {
#pragma unroll
            for (int p = 0; p < NM; ++p) uh[p] = ld_stream(src + t * NM + p);
        } else {
#pragma unroll
            for (int p = 0; p < NM; ++p) uh[p] = T(0);
        }
#pragma unroll
        for (int i = 0; i < NQ; ++i) {
            T s = uh[0] * tab.B[i];
#pragma unroll
            for (int p = 1; p < NM; ++p) s = fma(uh[p], tab.B[p * NQ + i], s);
            buf0[t * NQ + i] = s;  // s1[q][i]
        }
    }
    __syncthreads();
    if (active) {
        T s1[NM];
#pragma unroll
        for (int q = 0; q < NM; ++q) s1[q] = buf0[q * NQ + t];
#pragma unroll
        for (int j = 0; j < NQ; ++j) {
            T s = s1[0] * tab.B[j];
#pragma unroll
            for (int q = 1; q < NM; ++q) s = fma(s1[q], tab.B[q * NQ + j], s);
            u[j] = s;
        }
    } else {
#pragma unroll
        for (int j = 0; j < NQ; ++j) u[j] = T(0);
    }
}

// acc (registers, j-line of thread i) -> out = (B x B)^T acc.  Uses buf0; the caller
// must have synchronised any earlier readers of buf0.
template <int NM, int NQ, typename T>
__device__ __forceinline__ void quad_lines_to_modes(const Tab<NM, NQ, T> &tab,
                                                    const T (&acc)[NQ], T *__restrict__ dst,
                                                    bool valid, bool active, int t, T *buf0)
{
    if (active) {
#pragma unroll
        for (int q = 0; q < NM; ++q) {
            T s = acc[0] * tab.B[q * NQ];
#pragma unroll
            for (int j = 1; j < NQ; ++j) s = fma(acc[j], tab.B[q * NQ + j], s);
            buf0[q * NQ + t] = s;  // z[q][i]
        }
    }
    __syncthreads();
    if (active && t < NM && valid) {
        T v[NQ];
#pragma unroll
        for (int i = 0; i < NQ; ++i) v[i] = buf0[t * NQ + i];
#pragma unroll
        for (int p = 0; p < NM; ++p) {
            T s = v[0] * tab.B[p * NQ];
#pragma unroll
            for (int i = 1; i < NQ; ++i) s = fma(v[i], tab.B[p * NQ + i], s);
            st_stream(dst + t * NM + p, s);
        }
    }
}

template <int NM, int NQ, typename T, int DIR, int GEO>
__global__ __launch_bounds__(QuadCfg<NQ>::BLOCK) void quad_bwdtrans_kernel(
    const Tab<NM, NQ, T> tab, int nElmt, const T *__restrict__ in,
    const T *__restrict__ geom, T *__restrict__ out)
{
    using C = QuadCfg<NQ>;
    constexpr int M = NM * NM, Q = NQ * NQ;
    __shared__ T lds[C::EPB][Q];
    const int eb = threadIdx.x / C::TPE;
    const int t = threadIdx.x - eb * C::TPE;
    const bool active = eb < C::EPB;
    const int es = active ? eb : 0;
    const long e = (long)blockIdx.x * C::EPB + eb;
    const bool valid = active && e < nElmt;

    T u[NQ];
    quad_bwd_to_lines<NM, NQ, T>(tab, in + e * M, valid, active, t, lds[es], u);
    if (!valid) return;
    if (DIR == kTranspose) {
        if (GEO == kDeformed) {
#pragma unroll
            for (int j = 0; j < NQ; ++j) u[j] *= ld_stream(geom + e * Q + j * NQ + t);
        } else {
            const T wi = tab.w[t] * geom[e];
#pragma unroll
            for (int j = 0; j < NQ; ++j) u[j] *= wi * tab.w[j];
        }
    }
#pragma unroll
    for (int j = 0; j < NQ; ++j) st_stream(out + e * Q + j * NQ + t, u[j]);
}

template <int NM, int NQ, typename T, int GEO>
__global__ __launch_bounds__(QuadCfg<NQ>::BLOCK) void quad_iproduct_kernel(
    const Tab<NM, NQ, T> tab, int nElmt, const T *__restrict__ in,
    const T *__restrict__ geom, T *__restrict__ out)
{
    using C = QuadCfg<NQ>;
    constexpr int M = NM * NM, Q = NQ * NQ;
    __shared__ T lds[C::EPB][Q];
    const int eb = threadIdx.x / C::TPE;
    const int t = threadIdx.x - eb * C::TPE;
    const bool active = eb < C::EPB;
    const int es = active ? eb : 0;
    const long e = (long)blockIdx.x * C::EPB + eb;
    const bool valid = active && e < nElmt;

    T acc[NQ];
    if (valid) {
        if (GEO == kDeformed) {
#pragma unroll
            for (int j = 0; j < NQ; ++j)
                acc[j] = ld_stream(in + e * Q + j * NQ + t) * ld_stream(geom + e * Q + j * NQ + t);
        } else if (GEO == kRegular) {
            const T wi = tab.w[t] * geom[e];
#pragma unroll
            for (int j = 0; j < NQ; ++j) acc[j] = ld_stream(in + e * Q + j * NQ + t) * (wi * tab.w[j]);
        } else {
#pragma unroll
            for (int j = 0; j < NQ; ++j) acc[j] = ld_stream(in + e * Q + j * NQ + t);
        }
    } else {
#pragma unroll
        for (int j = 0; j < NQ; ++j) acc[j] = T(0);
    }
    quad_lines_to_modes<NM, NQ, T>(tab, acc, out + e * M, valid, active, t, lds[es]);
}

// Packed geometry: deformed [e][4][Q] (Jw, JwG00, JwG01, JwG11); regular [e][4]
// (J, J G00, J G01, J G11) with the weights applied here.
template <int NM, int NQ, typename T, int GEO>
__global__ __launch_bounds__(QuadCfg<NQ>::BLOCK) void quad_helmholtz_kernel(
    const Tab<NM, NQ, T> tab, int nElmt, T lambda, const T *__restrict__ in,
    const T *__restrict__ geom, T *__restrict__ out)
{
    using C = QuadCfg<NQ>;
    constexpr int M = NM * NM, Q = NQ * NQ;
    __shared__ T lds[C::EPB][2][Q];
    const int eb = threadIdx.x / C::TPE;
    const int t = threadIdx.x - eb * C::TPE;
    const bool active = eb < C::EPB;
    const int es = active ? eb : 0;
    const long e = (long)blockIdx.x * C::EPB + eb;
    const bool valid = active && e < nElmt;
    T *buf0 = lds[es][0], *buf1 = lds[es][1];

    T u[NQ];
    quad_bwd_to_lines<NM, NQ, T>(tab, in + e * M, valid, active, t, buf0, u);
    if (active) {
#pragma unroll
        for (int j = 0; j < NQ; ++j) buf1[j * NQ + t] = u[j];
    }
    __syncthreads();
    // d0 on i-lines of thread j = t
    if (active) {
        T v[NQ];
#pragma unroll
        for (int l = 0; l < NQ; ++l) v[l] = buf1[t * NQ + l];
#pragma unroll
        for (int i = 0; i < NQ; ++i) {
            T s = v[0] * tab.D[i];
#pragma unroll
            for (int l = 1; l < NQ; ++l) s = fma(v[l], tab.D[l * NQ + i], s);
            buf0[t * NQ + i] = s;
        }
    }
    __syncthreads();
    T g0[NQ], g1[NQ], acc[NQ];
    if (valid) {
        const T wi = tab.w[t];
#pragma unroll
        for (int j = 0; j < NQ; ++j) {
            T d1 = u[0] * tab.D[j];
#pragma unroll
            for (int l = 1; l < NQ; ++l) d1 = fma(u[l], tab.D[l * NQ + j], d1);
            const T d0 = buf0[j * NQ + t];
            T jw, m00, m01, m11;
            if (GEO == kDeformed) {
                const T *g = geom + e * (4 * Q) + j * NQ + t;
                jw = ld_stream(g);
                m00 = ld_stream(g + Q);
                m01 = ld_stream(g + 2 * Q);
                m11 = ld_stream(g + 3 * Q);
            } else {
                const T wq = wi * tab.w[j];
                const T *g = geom + e * 4;
                jw = g[0] * wq;
                m00 = g[1] * wq;
                m01 = g[2] * wq;
                m11 = g[3] * wq;
            }
            g0[j] = fma(m01, d1, m00 * d0);
            g1[j] = fma(m11, d1, m01 * d0);
            acc[j] = lambda * jw * u[j];
        }
    } else {
#pragma unroll
        for (int j = 0; j < NQ; ++j) g0[j] = g1[j] = acc[j] = T(0);
    }
    if (active) {
#pragma unroll
        for (int j = 0; j < NQ; ++j) buf1[j * NQ + t] = g0[j];
    }
#pragma unroll
    for (int j = 0; j < NQ; ++j) {
        T s = acc[j];
#pragma unroll
        for (int l = 0; l < NQ; ++l) s = fma(g1[l], tab.D[j * NQ + l], s);
        acc[j] = s;
    }
    __syncthreads();
    if (active) {
        T v[NQ];
#pragma unroll
        for (int l = 0; l < NQ; ++l) v[l] = buf1[t * NQ + l];
#pragma unroll
        for (int i = 0; i < NQ; ++i) {
            T s = v[0] * tab.D[i * NQ];
#pragma unroll
            for (int l = 1; l < NQ; ++l) s = fma(v[l], tab.D[i * NQ + l], s);
            buf0[t * NQ + i] = s;
        }
    }
    __syncthreads();
    if (active) {
#pragma unroll
        for (int j = 0; j < NQ; ++j) acc[j] += buf0[j * NQ + t];
    }
    __syncthreads();
    quad_lines_to_modes<NM, NQ, T>(tab, acc, out + e * M, valid, active, t, buf1);
}

// PhysDeriv, quads (PhysDeriv.h:260-400, PhysDerivKernels.hpp:93-137): out_0 = df0 d0 +
// df1 d1, out_1 = df2 d0 + df3 d1 (2D coordinates); layouts as the hex kernel.
template <int NQ, typename T, int GEO>
__global__ __launch_bounds__(QuadCfg<NQ>::BLOCK) void quad_physderiv_kernel(
    const DTab<NQ, T> tab, int nElmt, const T *__restrict__ in, const T *__restrict__ df,
    T *__restrict__ out)
{
    using C = QuadCfg<NQ>;
    constexpr int Q = NQ * NQ;
    __shared__ T lds[C::EPB][2][Q];
    const int eb = threadIdx.x / C::TPE;
    const int t = threadIdx.x - eb * C::TPE;
    const bool active = eb < C::EPB;
    const int es = active ? eb : 0;
    const long e = (long)blockIdx.x * C::EPB + eb;
    const bool valid = active && e < nElmt;
    T *bu = lds[es][0], *b0 = lds[es][1];

    T u[NQ];
#pragma unroll
    for (int j = 0; j < NQ; ++j) u[j] = valid ? ld_stream(in + e * Q + j * NQ + t) : T(0);
    if (active) {
#pragma unroll
        for (int j = 0; j < NQ; ++j) bu[j * NQ + t] = u[j];
    }
    __syncthreads();
    // d0 on i-lines of thread j = t
    if (active) {
        T v[NQ];
#pragma unroll
        for (int l = 0; l < NQ; ++l) v[l] = bu[t * NQ + l];
#pragma unroll
        for (int i = 0; i < NQ; ++i) {
            T s = v[0] * tab.D[i];
#pragma unroll
            for (int l = 1; l < NQ; ++l) s = fma(v[l], tab.D[l * NQ + i], s);
            b0[t * NQ + i] = s;
        }
    }
    __syncthreads();
    if (!valid) return;
    const long nt = (long)nElmt * Q;
#pragma unroll
    for (int j = 0; j < NQ; ++j) {
        T d1 = u[0] * tab.D[j];
#pragma unroll
        for (int l = 1; l < NQ; ++l) d1 = fma(u[l], tab.D[l * NQ + j], d1);
        const T d0 = b0[j * NQ + t];
        T f[4];
        if (GEO == kDeformed) {
#pragma unroll
            for (int d = 0; d < 4; ++d) f[d] = ld_stream(df + d * nt + e * Q + j * NQ + t);
        } else {
#pragma unroll
            for (int d = 0; d < 4; ++d) f[d] = df[(long)d * nElmt + e];
        }
        st_stream(out + e * Q + j * NQ + t, fma(d1, f[1], d0 * f[0]));
        st_stream(out + nt + e * Q + j * NQ + t, fma(d1, f[3], d0 * f[2]));
    }
}

// IProductWRTDerivBase, quads (IProductWRTDerivBase.h:371-, 2D coordinates): out =
// (D_0 B)^T diag(wJ) (df0 in0 + df2 in1) + (D_1 B)^T diag(wJ) (df1 in0 + df3 in1).
// Packed geometry: deformed [e][4][Q] = w J df[d]; regular [e][4] = J df[d].
template <int NM, int NQ, typename T, int GEO>
__global__ __launch_bounds__(QuadCfg<NQ>::BLOCK) void quad_iprod_deriv_kernel(
    const Tab<NM, NQ, T> tab, int nElmt, const T *__restrict__ in,
    const T *__restrict__ geom, T *__restrict__ out)
{
    using C = QuadCfg<NQ>;
    constexpr int M = NM * NM, Q = NQ * NQ;
    __shared__ T lds[C::EPB][2][Q];
    const int eb = threadIdx.x / C::TPE;
    const int t = threadIdx.x - eb * C::TPE;
    const bool active = eb < C::EPB;
    const int es = active ? eb : 0;
    const long e = (long)blockIdx.x * C::EPB + eb;
    const bool valid = active && e < nElmt;
    T *buf0 = lds[es][0], *buf1 = lds[es][1];
    const long nt = (long)nElmt * Q;

    T g0[NQ], g1[NQ];
    if (valid) {
        const T wi = tab.w[t];
#pragma unroll
        for (int j = 0; j < NQ; ++j) {
            const long o = e * Q + j * NQ + t;
            const T x0 = ld_stream(in + o), x1 = ld_stream(in + nt + o);
            T f[4];
            if (GEO == kDeformed) {
#pragma unroll
                for (int d = 0; d < 4; ++d) f[d] = ld_stream(geom + e * (4 * Q) + d * Q + j * NQ + t);
            } else {
                const T wq = wi * tab.w[j];
#pragma unroll
                for (int d = 0; d < 4; ++d) f[d] = geom[e * 4 + d] * wq;
            }
            g0[j] = fma(f[2], x1, f[0] * x0);
            g1[j] = fma(f[3], x1, f[1] * x0);
        }
    } else {
#pragma unroll
        for (int j = 0; j < NQ; ++j) g0[j] = g1[j] = T(0);
    }
    if (active) {
#pragma unroll
        for (int j = 0; j < NQ; ++j) buf1[j * NQ + t] = g0[j];
    }
    T acc[NQ];
#pragma unroll
    for (int j = 0; j < NQ; ++j) {
        T s = g1[0] * tab.D[j * NQ];
#pragma unroll
        for (int l = 1; l < NQ; ++l) s = fma(g1[l], tab.D[j * NQ + l], s);
        acc[j] = s;
    }
    __syncthreads();
    if (active) {
        T v[NQ];
#pragma unroll
        for (int l = 0; l < NQ; ++l) v[l] = buf1[t * NQ + l];
#pragma unroll
        for (int i = 0; i < NQ; ++i) {
            T s = v[0] * tab.D[i * NQ];
#pragma unroll
            for (int l = 1; l < NQ; ++l) s = fma(v[l], tab.D[i * NQ + l], s);
            buf0[t * NQ + i] = s;
        }
    }
    __syncthreads();
    if (active) {
#pragma unroll
        for (int j = 0; j < NQ; ++j) acc[j] += buf0[j * NQ + t];
    }
    __syncthreads();
    quad_lines_to_modes<NM, NQ, T>(tab, acc, out + e * M, valid, active, t, buf1);
}

}  // namespace imf
