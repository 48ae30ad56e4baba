// Geometry kernels: packing of (jac, df) into the streamed per-point form, and the
// geometric factors of straight-sided elements computed on the GPU.
#include "mf_geom.hpp"

#include <cmath>

namespace imf {

struct Pts {
    double v[16];
};

// Metric G^{ab} = sum_c df[c*dim+a] df[c*dim+b], accumulated in the reference order
// (Helmholtz.h:858-966 for hexes, :172-190 for quads).
__device__ __forceinline__ void metric3(const double *f, double *m)
{
    m[0] = f[0] * f[0]; m[0] += f[3] * f[3]; m[0] += f[6] * f[6];
    m[1] = f[0] * f[1]; m[1] += f[3] * f[4]; m[1] += f[6] * f[7];
    m[2] = f[0] * f[2]; m[2] += f[3] * f[5]; m[2] += f[6] * f[8];
    m[3] = f[1] * f[1]; m[3] += f[4] * f[4]; m[3] += f[7] * f[7];
    m[4] = f[1] * f[2]; m[4] += f[4] * f[5]; m[4] += f[7] * f[8];
    m[5] = f[2] * f[2]; m[5] += f[5] * f[5]; m[5] += f[8] * f[8];
}

__device__ __forceinline__ void metric2(const double *f, double *m)
{
    m[0] = f[0] * f[0]; m[0] += f[2] * f[2];
    m[1] = f[0] * f[1]; m[1] += f[2] * f[3];
    m[2] = f[1] * f[1]; m[2] += f[3] * f[3];
}

// Helmholtz, deformed: out[e][c][q] = Jw, Jw*G (6 or 3 unique entries), Jw = J w_i w_j (w_k).
template <typename T, int DIM>
__global__ void pack_helm_deformed(long nPts, int Q, int nq, Pts w, const double *__restrict__ jac,
                                   const double *__restrict__ df, T *__restrict__ out)
{
    constexpr int NC = DIM == 3 ? 7 : 4, NDF = DIM * DIM;
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= nPts) return;
    const long e = g / Q;
    const int q = (int)(g - e * Q);
    const int i = q % nq, j = (q / nq) % nq, k = q / (nq * nq);
    double wq = w.v[i] * w.v[j];
    if (DIM == 3) wq *= w.v[k];
    const double jw = jac[g] * wq;
    double f[NDF], m[6];
#pragma unroll
    for (int d = 0; d < NDF; ++d) f[d] = df[(long)d * nPts + g];
    if (DIM == 3) metric3(f, m); else metric2(f, m);
    // hexes: per-element blocks padded to whole 128-byte lines (geo_stride)
    const long stride = DIM == 3 ? geo_stride((long)NC * Q, sizeof(T)) : (long)NC * Q;
    T *o = out + e * stride + q;
    o[0] = T(jw);
#pragma unroll
    for (int c = 1; c < NC; ++c) o[(long)c * Q] = T(jw * m[c - 1]);
}

// Helmholtz, regular: out[e][c] = J, J*G (weights applied by the operator kernel).
template <typename T, int DIM>
__global__ void pack_helm_regular(long nElmt, const double *__restrict__ jac,
                                  const double *__restrict__ df, T *__restrict__ out)
{
    constexpr int NC = DIM == 3 ? 7 : 4, NDF = DIM * DIM;
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nElmt) return;
    double f[NDF], m[6];
#pragma unroll
    for (int d = 0; d < NDF; ++d) f[d] = df[(long)d * nElmt + e];
    if (DIM == 3) metric3(f, m); else metric2(f, m);
    const double J = jac[e];
    out[e * NC] = T(J);
#pragma unroll
    for (int c = 1; c < NC; ++c) out[e * NC + c] = T(J * m[c - 1]);
}

// IProduct: deformed out[e][q] = J w_i w_j (w_k); regular out[e] = J.
template <typename T, int DIM>
__global__ void pack_jw(long n, int Q, int nq, int deformed, Pts w,
                        const double *__restrict__ jac, T *__restrict__ out)
{
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    if (!deformed) {
        out[g] = T(jac[g]);
        return;
    }
    const int q = (int)(g % Q);
    const int i = q % nq, j = (q / nq) % nq, k = q / (nq * nq);
    double wq = w.v[i] * w.v[j];
    if (DIM == 3) wq *= w.v[k];
    out[g] = T(jac[g] * wq);
}

// PhysDeriv: df converted to the operator's precision, layout unchanged ([d][n]).
template <typename T>
__global__ void pack_df(long n, const double *__restrict__ df, T *__restrict__ out)
{
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g < n) out[g] = T(df[g]);
}

// IProductWRTDerivBase, deformed: out[e][d][q] = J w_i w_j (w_k) df[d] (hex element blocks
// padded to whole 128-byte lines); regular: out[e][d] = J df[d].
template <typename T, int DIM>
__global__ void pack_ipd_deformed(long nPts, int Q, int nq, Pts w,
                                  const double *__restrict__ jac,
                                  const double *__restrict__ df, T *__restrict__ out)
{
    constexpr int NDF = DIM * DIM;
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= nPts) return;
    const long e = g / Q;
    const int q = (int)(g - e * Q);
    const int i = q % nq, j = (q / nq) % nq, k = q / (nq * nq);
    double wq = w.v[i] * w.v[j];
    if (DIM == 3) wq *= w.v[k];
    const double jw = jac[g] * wq;
    const long stride = DIM == 3 ? geo_stride((long)NDF * Q, sizeof(T)) : (long)NDF * Q;
    T *o = out + e * stride + q;
#pragma unroll
    for (int d = 0; d < NDF; ++d) o[(long)d * Q] = T(jw * df[(long)d * nPts + g]);
}

template <typename T, int DIM>
__global__ void pack_ipd_regular(long nElmt, const double *__restrict__ jac,
                                 const double *__restrict__ df, T *__restrict__ out)
{
    constexpr int NDF = DIM * DIM;
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nElmt) return;
#pragma unroll
    for (int d = 0; d < NDF; ++d) out[e * NDF + d] = T(jac[e] * df[(long)d * nElmt + e]);
}

hipError_t pack_geometry(int kind, int shape, int deformed, int prec, long nElmt, int nq,
                         const double *w, const double *d_jac, const double *d_df,
                         void *d_out, hipStream_t s)
{
    Pts pw{};
    for (int i = 0; i < nq && i < 16; ++i) pw.v[i] = w[i];
    const int Q = shape == kHex ? nq * nq * nq : nq * nq;
    const int bs = 256;
    const int ndf = shape == kHex ? 9 : 4;
    if (kind == kPhysDeriv) {
        const long n = (long)ndf * (deformed ? nElmt * Q : nElmt);
        const dim3 g((unsigned)((n + bs - 1) / bs));
        if (prec == 0) pack_df<double><<<g, bs, 0, s>>>(n, d_df, (double *)d_out);
        else pack_df<float><<<g, bs, 0, s>>>(n, d_df, (float *)d_out);
        return hipGetLastError();
    }
    if (kind == kIProdDeriv) {
        if (deformed) {
            const long n = nElmt * Q;
            const dim3 g((unsigned)((n + bs - 1) / bs));
            if (shape == kHex) {
                if (prec == 0) pack_ipd_deformed<double, 3><<<g, bs, 0, s>>>(n, Q, nq, pw, d_jac, d_df, (double *)d_out);
                else pack_ipd_deformed<float, 3><<<g, bs, 0, s>>>(n, Q, nq, pw, d_jac, d_df, (float *)d_out);
            } else {
                if (prec == 0) pack_ipd_deformed<double, 2><<<g, bs, 0, s>>>(n, Q, nq, pw, d_jac, d_df, (double *)d_out);
                else pack_ipd_deformed<float, 2><<<g, bs, 0, s>>>(n, Q, nq, pw, d_jac, d_df, (float *)d_out);
            }
        } else {
            const dim3 g((unsigned)((nElmt + bs - 1) / bs));
            if (shape == kHex) {
                if (prec == 0) pack_ipd_regular<double, 3><<<g, bs, 0, s>>>(nElmt, d_jac, d_df, (double *)d_out);
                else pack_ipd_regular<float, 3><<<g, bs, 0, s>>>(nElmt, d_jac, d_df, (float *)d_out);
            } else {
                if (prec == 0) pack_ipd_regular<double, 2><<<g, bs, 0, s>>>(nElmt, d_jac, d_df, (double *)d_out);
                else pack_ipd_regular<float, 2><<<g, bs, 0, s>>>(nElmt, d_jac, d_df, (float *)d_out);
            }
        }
        return hipGetLastError();
    }
    if (kind == kHelmholtz) {
        if (deformed) {
            const long n = nElmt * Q;
            const dim3 g((unsigned)((n + bs - 1) / bs));
            if (shape == kHex) {
                if (prec == 0) pack_helm_deformed<double, 3><<<g, bs, 0, s>>>(n, Q, nq, pw, d_jac, d_df, (double *)d_out);
                else pack_helm_deformed<float, 3><<<g, bs, 0, s>>>(n, Q, nq, pw, d_jac, d_df, (float *)d_out);
            } else {
                if (prec == 0) pack_helm_deformed<double, 2><<<g, bs, 0, s>>>(n, Q, nq, pw, d_jac, d_df, (double *)d_out);
                else pack_helm_deformed<float, 2><<<g, bs, 0, s>>>(n, Q, nq, pw, d_jac, d_df, (float *)d_out);
            }
        } else {
            const dim3 g((unsigned)((nElmt + bs - 1) / bs));
            if (shape == kHex) {
                if (prec == 0) pack_helm_regular<double, 3><<<g, bs, 0, s>>>(nElmt, d_jac, d_df, (double *)d_out);
                else pack_helm_regular<float, 3><<<g, bs, 0, s>>>(nElmt, d_jac, d_df, (float *)d_out);
            } else {
                if (prec == 0) pack_helm_regular<double, 2><<<g, bs, 0, s>>>(nElmt, d_jac, d_df, (double *)d_out);
                else pack_helm_regular<float, 2><<<g, bs, 0, s>>>(nElmt, d_jac, d_df, (float *)d_out);
            }
        }
    } else {
        const long n = deformed ? nElmt * Q : nElmt;
        const dim3 g((unsigned)((n + bs - 1) / bs));
        if (shape == kHex) {
            if (prec == 0) pack_jw<double, 3><<<g, bs, 0, s>>>(n, Q, nq, deformed, pw, d_jac, (double *)d_out);
            else pack_jw<float, 3><<<g, bs, 0, s>>>(n, Q, nq, deformed, pw, d_jac, (float *)d_out);
        } else {
            if (prec == 0) pack_jw<double, 2><<<g, bs, 0, s>>>(n, Q, nq, deformed, pw, d_jac, (double *)d_out);
            else pack_jw<float, 2><<<g, bs, 0, s>>>(n, Q, nq, deformed, pw, d_jac, (float *)d_out);
        }
    }
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Geometric factors of straight-sided elements (GeomFactors.cpp:399-478).
// Map derivatives deriv[i][k] = dx_k/dxi_i of the trilinear / bilinear vertex map
// (vertex order of StdHexExp.cpp:1042-1082 / StdQuadExp), then g = deriv deriv^T,
// jac = sqrt(det g), df[k*dim+j] = sum_i deriv[i][k] (g^{-1})[i][j].
// ---------------------------------------------------------------------------

__device__ void factors_from_deriv(int dim, const double *deriv, double *jac, double *fac)
{
    double g[9], gi[9];
    for (int i = 0, l = 0; i < dim; ++i)
        for (int j = 0; j < dim; ++j, ++l) {
            double s = 0.0;
            for (int k = 0; k < dim; ++k) s += deriv[i * dim + k] * deriv[j * dim + k];
            g[l] = s;
        }
    double det = 0.0;
    if (dim == 3) {
        gi[0] = g[4] * g[8] - g[5] * g[7];
        gi[1] = g[2] * g[7] - g[1] * g[8];
        gi[2] = g[1] * g[5] - g[2] * g[4];
        gi[3] = g[5] * g[6] - g[3] * g[8];
        gi[4] = g[0] * g[8] - g[2] * g[6];
        gi[5] = g[2] * g[3] - g[0] * g[5];
        gi[6] = g[3] * g[7] - g[4] * g[6];
        gi[7] = g[1] * g[6] - g[0] * g[7];
        gi[8] = g[0] * g[4] - g[1] * g[3];
        for (int i = 0; i < 3; ++i) det += g[i] * gi[i * 3];
    } else {
        gi[0] = g[3]; gi[1] = -g[1]; gi[2] = -g[2]; gi[3] = g[0];
        for (int i = 0; i < 2; ++i) det += g[i] * gi[i * 2];
    }
    for (int i = 0; i < dim * dim; ++i) gi[i] /= det;
    *jac = sqrt(det);
    for (int k = 0, l = 0; k < dim; ++k)
        for (int j = 0; j < dim; ++j, ++l) {
            double s = 0.0;
            for (int i = 0; i < dim; ++i) s += deriv[i * dim + k] * gi[dim * i + j];
            fac[l] = s;
        }
}

// Points (z0, z1, z2) per direction, nq0 x nq1 x nq2 per element (all equal for the
// isotropic operators).
__global__ void geom_hex_kernel(const double *__restrict__ verts, long nElmt, int nq0, int nq1,
                                int nq2, Pts z0, Pts z1, Pts z2, int deformed,
                                double *__restrict__ jac, double *__restrict__ df)
{
    const int Q = deformed ? nq0 * nq1 * nq2 : 1;
    const long n = nElmt * Q;
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    const long e = g / Q;
    const int q = (int)(g - e * Q);
    const int i = q % nq0, j = (q / nq0) % nq1, k = q / (nq0 * nq1);
    const double xi[3] = {z0.v[i], z1.v[j], z2.v[k]};
    const double *v = verts + e * 24;
    double deriv[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int nv = 0; nv < 8; ++nv) {
        const int c[3] = {(nv % 4) % 3 > 0, (nv % 4) > 1, nv > 3};
        double f[3], d[3];
        for (int a = 0; a < 3; ++a) {
            f[a] = c[a] ? 0.5 * (1 + xi[a]) : 0.5 * (1 - xi[a]);
            d[a] = c[a] ? 0.5 : -0.5;
        }
        const double dN[3] = {d[0] * f[1] * f[2], f[0] * d[1] * f[2], f[0] * f[1] * d[2]};
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) deriv[a * 3 + b] += dN[a] * v[nv * 3 + b];
    }
    double J, fac[9];
    factors_from_deriv(3, deriv, &J, fac);
    jac[g] = J;
    for (int d = 0; d < 9; ++d) df[(long)d * n + g] = fac[d];
}

__global__ void geom_quad_kernel(const double *__restrict__ verts, long nElmt, int nq, Pts z,
                                 int deformed, double *__restrict__ jac, double *__restrict__ df)
{
    const int Q = deformed ? nq * nq : 1;
    const long n = nElmt * Q;
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    const long e = g / Q;
    const int q = (int)(g - e * Q);
    const double xi[2] = {z.v[q % nq], z.v[q / nq]};
    const double *v = verts + e * 8;
    double deriv[4] = {0, 0, 0, 0};
    for (int nv = 0; nv < 4; ++nv) {
        const int c[2] = {nv % 3 > 0, nv > 1};
        double f[2], d[2];
        for (int a = 0; a < 2; ++a) {
            f[a] = c[a] ? 0.5 * (1 + xi[a]) : 0.5 * (1 - xi[a]);
            d[a] = c[a] ? 0.5 : -0.5;
        }
        const double dN[2] = {d[0] * f[1], f[0] * d[1]};
        for (int a = 0; a < 2; ++a)
            for (int b = 0; b < 2; ++b) deriv[a * 2 + b] += dN[a] * v[nv * 2 + b];
    }
    double J, fac[4];
    factors_from_deriv(2, deriv, &J, fac);
    jac[g] = J;
    for (int d = 0; d < 4; ++d) df[(long)d * n + g] = fac[d];
}

// Straight-sided triangles (TriGeom: x(xi) = -(xi0 + xi1)/2 v0 + (1 + xi0)/2 v1 +
// (1 + xi1)/2 v2): a constant map derivative, so Q copies per element when deformed.
__global__ void geom_tri_kernel(const double *__restrict__ verts, long nElmt, int Q, int deformed,
                                double *__restrict__ jac, double *__restrict__ df)
{
    const int per = deformed ? Q : 1;
    const long n = nElmt * per;
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    const long e = g / per;
    const double *v = verts + e * 6;
    double deriv[4];
    for (int c = 0; c < 2; ++c) {
        deriv[0 * 2 + c] = 0.5 * (v[2 + c] - v[c]);
        deriv[1 * 2 + c] = 0.5 * (v[4 + c] - v[c]);
    }
    double J, fac[4];
    factors_from_deriv(2, deriv, &J, fac);
    jac[g] = J;
    for (int d = 0; d < 4; ++d) df[(long)d * n + g] = fac[d];
}

hipError_t geom_tri(const double *d_verts, long nElmt, int Q, int deformed, double *d_jac,
                    double *d_df, hipStream_t s)
{
    const long n = nElmt * (deformed ? Q : 1);
    const int bs = 256;
    geom_tri_kernel<<<(unsigned)((n + bs - 1) / bs), bs, 0, s>>>(d_verts, nElmt, Q, deformed,
                                                                 d_jac, d_df);
    return hipGetLastError();
}

// Straight-sided tetrahedra, vertices in StdTetExp order (xi = (-1,-1,-1), (1,-1,-1),
// (-1,1,-1), (-1,-1,1)): x = -(1 + xi0 + xi1 + xi2)/2 v0 + sum_a (1 + xi_a)/2 v_{a+1}, an
// affine map, so the factors are constant (repeated at the Q points when deformed).
__global__ void geom_tet_kernel(const double *__restrict__ verts, long nElmt, int Q, int deformed,
                                double *__restrict__ jac, double *__restrict__ df)
{
    const int per = deformed ? Q : 1;
    const long n = nElmt * per;
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    const long e = g / per;
    const double *v = verts + e * 12;
    double deriv[9];
    for (int a = 0; a < 3; ++a)
        for (int c = 0; c < 3; ++c) deriv[a * 3 + c] = 0.5 * (v[(a + 1) * 3 + c] - v[c]);
    double J, fac[9];
    factors_from_deriv(3, deriv, &J, fac);
    jac[g] = J;
    for (int d = 0; d < 9; ++d) df[(long)d * n + g] = fac[d];
}

hipError_t geom_tet(const double *d_verts, long nElmt, int Q, int deformed, double *d_jac,
                    double *d_df, hipStream_t s)
{
    const long n = nElmt * (deformed ? Q : 1);
    const int bs = 256;
    geom_tet_kernel<<<(unsigned)((n + bs - 1) / bs), bs, 0, s>>>(d_verts, nElmt, Q, deformed,
                                                                 d_jac, d_df);
    return hipGetLastError();
}

// Straight-sided pyramids, vertices in StdPyrExp order (xi = (-1,-1,-1), (1,-1,-1),
// (1,1,-1), (-1,1,-1), apex (-1,-1,1)): x = sum of the vertex modes a0 b0 c0, a1 b0 c0,
// a1 b1 c0, a0 b1 c0, c1 (a = (1 -/+ eta0)/2, b the same in eta1, c in eta2) times the
// vertices, differentiated in eta and turned into xi derivatives by the collapsed chain
// rule d/dxi0 = h2 d/deta0, d/dxi1 = h2 d/deta1, d/dxi2 = h0 h2 d/deta0 + h1 h2 d/deta1 +
// d/deta2.  A pyramid whose base is not a parallelogram has a non-affine map, so deformed
// factors vary over the points; regular ones are taken at point 0.
__global__ void geom_pyr_kernel(const double *__restrict__ verts, long nElmt, int nq0, int nq1,
                                int nq2, Pts z0, Pts z1, Pts z2, int deformed,
                                double *__restrict__ jac, double *__restrict__ df)
{
    const int Q = deformed ? nq0 * nq1 * nq2 : 1;
    const long n = nElmt * Q;
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    const long e = g / Q;
    const int q = (int)(g - e * Q);
    const int i = q % nq0, j = (q / nq0) % nq1, k = q / (nq0 * nq1);
    const double e0 = z0.v[i], e1 = z1.v[j], e2 = z2.v[k];
    const double a0 = 0.5 * (1 - e0), a1 = 0.5 * (1 + e0), b0 = 0.5 * (1 - e1),
                 b1 = 0.5 * (1 + e1), c0 = 0.5 * (1 - e2);
    const double dN[5][3] = {{-0.5 * b0 * c0, -0.5 * a0 * c0, -0.5 * a0 * b0},
                             {0.5 * b0 * c0, -0.5 * a1 * c0, -0.5 * a1 * b0},
                             {0.5 * b1 * c0, 0.5 * a1 * c0, -0.5 * a1 * b1},
                             {-0.5 * b1 * c0, 0.5 * a0 * c0, -0.5 * a0 * b1},
                             {0.0, 0.0, 0.5}};
    const double h2 = 2.0 / (1.0 - e2);
    const double *v = verts + e * 15;
    double deriv[9];
    for (int c = 0; c < 3; ++c) {
        double de[3] = {0, 0, 0};
        for (int nv = 0; nv < 5; ++nv)
            for (int a = 0; a < 3; ++a) de[a] += dN[nv][a] * v[nv * 3 + c];
        deriv[0 * 3 + c] = h2 * de[0];
        deriv[1 * 3 + c] = h2 * de[1];
        deriv[2 * 3 + c] = a1 * h2 * de[0] + b1 * h2 * de[1] + de[2];
    }
    double J, fac[9];
    factors_from_deriv(3, deriv, &J, fac);
    jac[g] = J;
    for (int d = 0; d < 9; ++d) df[(long)d * n + g] = fac[d];
}

hipError_t geom_pyr(const double *d_verts, long nElmt, const int *nq, const double *const *z,
                    int deformed, double *d_jac, double *d_df, hipStream_t s)
{
    Pts pz[3] = {};
    for (int d = 0; d < 3; ++d)
        for (int i = 0; i < nq[d] && i < 16; ++i) pz[d].v[i] = z[d][i];
    const long n = nElmt * (deformed ? (long)nq[0] * nq[1] * nq[2] : 1);
    const int bs = 256;
    geom_pyr_kernel<<<(unsigned)((n + bs - 1) / bs), bs, 0, s>>>(
        d_verts, nElmt, nq[0], nq[1], nq[2], pz[0], pz[1], pz[2], deformed, d_jac, d_df);
    return hipGetLastError();
}

// Straight-sided prisms, vertices in StdPrismExp order: 0..3 the xi2 = -1 quad (xi0, xi1
// = (-1,-1), (1,-1), (1,1), (-1,1)), 4 and 5 the collapsed edge at xi2 = 1 (xi1 = -1, 1).
// The map is linear in each of (xi0, xi2) and xi1 but not affine overall, so deformed
// factors vary over the points; regular ones are taken at point 0.  Point (i, j, k) is at
// xi = ((1 + z0_i)(1 - z2_k)/2 - 1, z1_j, z2_k).
__global__ void geom_prism_kernel(const double *__restrict__ verts, long nElmt, int nq0,
                                  int nq1, int nq2, Pts z0, Pts z1, Pts z2, int deformed,
                                  double *__restrict__ jac, double *__restrict__ df)
{
    const int Q = deformed ? nq0 * nq1 * nq2 : 1;
    const long n = nElmt * Q;
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    const long e = g / Q;
    const int q = (int)(g - e * Q);
    const int i = q % nq0, j = (q / nq0) % nq1, k = q / (nq0 * nq1);
    const double x0 = 0.5 * (1.0 + z0.v[i]) * (1.0 - z2.v[k]) - 1.0, x1 = z1.v[j], x2 = z2.v[k];
    // triangle (xi0, xi2) shape functions times the segment ones of xi1
    const double la = -0.5 * (x0 + x2), lb = 0.5 * (1.0 + x0), lc = 0.5 * (1.0 + x2);
    const double sl = 0.5 * (1.0 - x1), sh = 0.5 * (1.0 + x1);
    const double dN[6][3] = {{-0.5 * sl, -0.5 * la, -0.5 * sl}, {0.5 * sl, -0.5 * lb, 0.0},
                             {0.5 * sh, 0.5 * lb, 0.0},         {-0.5 * sh, 0.5 * la, -0.5 * sh},
                             {0.0, -0.5 * lc, 0.5 * sl},        {0.0, 0.5 * lc, 0.5 * sh}};
    const double *v = verts + e * 18;
    double deriv[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int nv = 0; nv < 6; ++nv)
        for (int a = 0; a < 3; ++a)
            for (int c = 0; c < 3; ++c) deriv[a * 3 + c] += dN[nv][a] * v[nv * 3 + c];
    double J, fac[9];
    factors_from_deriv(3, deriv, &J, fac);
    jac[g] = J;
    for (int d = 0; d < 9; ++d) df[(long)d * n + g] = fac[d];
}

hipError_t geom_prism(const double *d_verts, long nElmt, const int *nq, const double *const *z,
                      int deformed, double *d_jac, double *d_df, hipStream_t s)
{
    Pts pz[3] = {};
    for (int d = 0; d < 3; ++d)
        for (int i = 0; i < nq[d] && i < 16; ++i) pz[d].v[i] = z[d][i];
    const long n = nElmt * (deformed ? (long)nq[0] * nq[1] * nq[2] : 1);
    const int bs = 256;
    geom_prism_kernel<<<(unsigned)((n + bs - 1) / bs), bs, 0, s>>>(
        d_verts, nElmt, nq[0], nq[1], nq[2], pz[0], pz[1], pz[2], deformed, d_jac, d_df);
    return hipGetLastError();
}

hipError_t geom_from_vertices(int shape, const double *d_verts, long nElmt, int nq,
                              const double *z, int deformed, double *d_jac, double *d_df,
                              hipStream_t s)
{
    Pts pz{};
    for (int i = 0; i < nq && i < 16; ++i) pz.v[i] = z[i];
    const long Q = deformed ? (shape == kHex ? (long)nq * nq * nq : (long)nq * nq) : 1;
    const long n = nElmt * Q;
    const int bs = 256;
    const dim3 g((unsigned)((n + bs - 1) / bs));
    if (shape == kHex)
        geom_hex_kernel<<<g, bs, 0, s>>>(d_verts, nElmt, nq, nq, nq, pz, pz, pz, deformed, d_jac,
                                         d_df);
    else
        geom_quad_kernel<<<g, bs, 0, s>>>(d_verts, nElmt, nq, pz, deformed, d_jac, d_df);
    return hipGetLastError();
}

hipError_t geom_hex_aniso(const double *d_verts, long nElmt, const int *nq,
                          const double *const *z, int deformed, double *d_jac, double *d_df,
                          hipStream_t s)
{
    Pts pz[3] = {};
    for (int d = 0; d < 3; ++d)
        for (int i = 0; i < nq[d] && i < 16; ++i) pz[d].v[i] = z[d][i];
    const long Q = deformed ? (long)nq[0] * nq[1] * nq[2] : 1;
    const long n = nElmt * Q;
    const int bs = 256;
    const dim3 g((unsigned)((n + bs - 1) / bs));
    geom_hex_kernel<<<g, bs, 0, s>>>(d_verts, nElmt, nq[0], nq[1], nq[2], pz[0], pz[1], pz[2],
                                     deformed, d_jac, d_df);
    return hipGetLastError();
}

}  // namespace imf
